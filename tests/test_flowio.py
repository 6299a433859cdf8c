"""Rocflu flow-file and restart-info formats (rocflupicl_amd/flowio.py) on
the CPU: Fortran edit descriptors against hand-worked strings, the section
layout of RFLU_WriteFlowASCII / RFLU_WriteFlowBinary, and round trips."""
import struct

import numpy as np

from rocflupicl_amd import flowio


def test_fortran_edit_descriptors():
    assert flowio.fortran_e23_16(1.0) == " 0.1000000000000000E+01"
    assert flowio.fortran_e23_16(-2.5e-7) == "-0.2500000000000000E-06"
    assert flowio.fortran_e23_16(123456.789) == " 0.1234567890000000E+06"
    assert flowio.fortran_e23_16(0.0) == " 0.0000000000000000E+00"
    assert flowio.fortran_e23_16(1e-300) == " 0.1000000000000000-299"
    assert all(len(flowio.fortran_e23_16(x)) == 23 for x in (1.0, -1e200, 3e-5))
    assert flowio.file_name("rectshktb", ".mixt.cva", 1, 1.5e-3) == "rectshktb.mixt.cva_00001_1.50000E-03"


def test_ascii_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    cv = rng.standard_normal((7, 5)) * np.array([1.0, 300.0, 300.0, 300.0, 2.5e5])
    p = tmp_path / "case.mixt.cva_00001_1.00000E-03"
    flowio.write_flow_ascii(p, cv, 1.0e-3, res_init=0.0)
    lines = p.read_text().splitlines()
    assert lines[:9] == ["# ROCFLU flow file (v2)", "# Precision and range", "      15     307",
                         "# Initial residual", " 0.0000000000000000E+00", "# Physical time",
                         " 0.1000000000000000E-02", "# Dimensions", "               7               5"]
    assert lines[9] == "#  1" and len(lines[10]) == 5 * 23 and len(lines[11]) == 2 * 23
    assert lines[-1] == "# End"
    d = flowio.read_flow_ascii(p)
    assert d["time"] == 1.0e-3
    assert np.allclose(d["cv"], cv, rtol=6e-16, atol=0)


def test_binary_records_and_exact_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    cv = rng.standard_normal((11, 5))
    p = tmp_path / "case.mixt.cv_00001_2.00000E-03"
    flowio.write_flow_binary(p, cv, 2.0e-3, res_init=0.5)
    raw = p.read_bytes()
    assert struct.unpack_from("<i", raw, 0)[0] == 80 and raw[4:27] == b"# ROCFLU flow file (v2)"
    d = flowio.read_flow_binary(p)
    assert d["time"] == 2.0e-3 and d["res_init"] == 0.5
    assert np.array_equal(d["cv"], cv)


def test_restart_info_appends(tmp_path):
    flowio.append_restart_info("case", 1.0e-3, tmp_path)
    flowio.append_restart_info("case", 0.25, tmp_path)
    assert flowio.read_restart_info("case", tmp_path) == [1.0e-3, 0.25]


def test_particle_vtu_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    n = 9
    y = rng.standard_normal((n, 12))
    rprop = rng.standard_normal((n, 48))
    tags = np.stack([np.zeros(n), np.arange(1, n + 1), np.arange(1, n + 1)], axis=1)
    p = tmp_path / "par00001.vtu"
    flowio.write_particle_vtu(p, y, rprop, tags, time=2.5e-4, cycle=7)
    raw = p.read_bytes()
    assert raw.startswith(b'<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian">\n')
    assert b'Name="TIME" NumberOfTuples="1" format="ascii">  0.2500000E-03 </DataArray>' in raw
    # the 67 arrays follow each other: offset = previous + 4 * components * n + 4
    assert b'Name="y01" NumberOfComponents="1" format="append" offset="112"/>' in raw
    assert raw.endswith(b"</AppendedData>\n</VTKFile>\n")
    d = flowio.read_particle_vtu(p)
    assert d["time"] == np.float64(np.float32(2.5e-4)) or abs(d["time"] - 2.5e-4) < 1e-10
    assert d["cycle"] == 7
    assert np.array_equal(d["y"], y.astype(np.float32).astype(np.float64))
    assert np.array_equal(d["rprop"], rprop.astype(np.float32).astype(np.float64))
    assert np.array_equal(d["position"], y[:, :3].astype(np.float32).astype(np.float64))
    assert np.array_equal(d["tags"], tags)


def test_dependent_variable_file_layout(tmp_path):
    """The .mixt.dva companion: header as the flow file, dimensions (cells, 4),
    p, T, c, then dv(4, j) read past each 3-row column (the next cell's p)."""
    rng = np.random.default_rng(6)
    n = 7
    dv = np.column_stack([1e5 * (1 + rng.random(n)), 300 * (1 + rng.random(n)), 340 * (1 + rng.random(n))])
    p = tmp_path / flowio.file_name("case", ".mixt.dva", 1, 2.5e-4)
    flowio.write_dv_ascii(p, dv, 2.5e-4)
    text = p.read_text().splitlines()
    assert text[0] == "# ROCFLU flow file (v2)"
    dims = text[text.index("# Dimensions") + 1]
    assert dims == f"{n:16d}{4:16d}"
    assert [ln for ln in text if ln.startswith("# ") and ln[2:].strip().isdigit()] == ["#  1", "#  2", "#  3", "#  4"]
    d = flowio.read_flow_ascii(p)
    assert d["cv"].shape == (n, 4) and d["time"] == 2.5e-4
    assert np.allclose(d["cv"][:, :3], dv, rtol=1e-15, atol=0)
    assert np.array_equal(d["cv"][:n - 1, 3], d["cv"][1:, 0]) and d["cv"][n - 1, 3] == 0.0
