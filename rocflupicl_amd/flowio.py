"""Rocflu flow and restart-info files (modflu/RFLU_ModReadWriteFlow.F90,
libflu/RFLU_WriteRestartInfo.F90), so that GPU runs can start from a
reference solution file and write files the reference (and its
post-processors) read back.

* ASCII flow file ``<case>.mixt.cva_<region:05d>_<time:1PE11.5>``
  (RFLU_WriteFlowASCII, :2187-2367; read by RFLU_ReadFlowASCII, :1078-1301):
  section lines ``# ROCFLU flow file (v2)``, ``# Precision and range``
  (2I8), ``# Initial residual`` (E23.16), ``# Physical time`` (E23.16),
  ``# Dimensions`` (2I16: cells, variables), then per conserved variable
  ``# <I2>`` and its values five per line in E23.16, and ``# End``.
* Binary flow file ``<case>.mixt.cv_...`` (RFLU_WriteFlowBinary,
  :2392-2559): the same sections as Fortran unformatted sequential records
  (4-byte length markers around each record; section strings are
  CHARACTER(80), CHRLEN in modfloflu/ModDataTypes.F90:73; integers 4 bytes).
  Little-endian, as the reference's FORMAT_BINARY and FORMAT_BINARY_L write
  on x86-64.
* ASCII dependent-variable file ``<case>.mixt.dva_...``, written beside
  every ASCII flow file (RFLU_WriteFlowASCII, :2228-2230, 2323-2352): the
  same header, dimensions (cells, 4) and four sections.  A perfect gas has
  three dependent variables (p, T, c; nDv = 3,
  libflu/RFLU_DerivedInputValues.F90:186-193), so the reference's fourth
  section reads dv(4, j) past the end of each 3-row column of its
  column-major array: dv(1, j+1), the next cell's pressure.  That is
  written here too; for the last cell the read lies beyond the array and 0
  is written.  The binary writer has no such companion.
* Restart info ``<case>.rin``: one line per written solution, the physical
  time, list-directed (RFLU_WriteRestartInfo.F90), appended.

The E23.16 text form keeps 16 significant digits, so an ASCII round trip may
change the last bit of a value; the binary form is exact.  Cells are in the
region's (reference) numbering; cv is (nCells, nCv) with the conserved
variables rho, rho*u, rho*v, rho*w, rho*E.
"""
import os
import struct

import numpy as np

CHRLEN = 80
PRECISION, RANGE = 15, 307   # PRECISION(1.0_RFREAL), RANGE(1.0_RFREAL) for real*8


def file_name(case, ext, region, time, directory=""):
    """BuildFileNameUnsteady (libfloflu/BuildFileNameUnsteady.F90):
    dir + case + ext + '_' + I5.5 + '_' + 1PE11.5."""
    return os.path.join(directory, f"{case}{ext}_{region:05d}_{fortran_1pe(time, 11, 5)}")


def fortran_1pe(x, w, d):
    """Fortran 1PEw.d: one digit before the point, d after, two-digit exponent."""
    s = f"{x:.{d}E}"
    mant, exp = s.split("E")
    e = int(exp)
    out = f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}"
    return out.rjust(w)


def fortran_e23_16(x):
    """Fortran E23.16: 0.dddddddddddddddd E+xx (a 3-digit exponent drops the
    'E', as gfortran prints it)."""
    x = float(x)
    if x == 0.0:
        return " 0.0000000000000000E+00"
    s = f"{x:.15e}"                        # d.ddddddddddddddde+XX, 16 significant digits
    mant, exp = s.split("e")
    sign = "-" if mant.startswith("-") else ""
    digits = mant.lstrip("-").replace(".", "")
    e = int(exp) + 1
    es = f"E{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) < 100 else f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    return f"{sign}0.{digits}{es}".rjust(23)


def _parse_e(tok):
    tok = tok.strip()
    if "E" not in tok.upper():  # 0.1234+100 form
        for k in range(len(tok) - 1, 0, -1):
            if tok[k] in "+-":
                tok = tok[:k] + "E" + tok[k:]
                break
    return float(tok)


def write_flow_ascii(path, cv, time, res_init=0.0):
    cv = np.asarray(cv, np.float64)
    n, nv = cv.shape
    lines = ["# ROCFLU flow file (v2)", "# Precision and range", f"{PRECISION:8d}{RANGE:8d}",
             "# Initial residual", fortran_e23_16(res_init), "# Physical time", fortran_e23_16(time),
             "# Dimensions", f"{n:16d}{nv:16d}"]
    for k in range(nv):
        lines.append(f"# {k + 1:2d}")
        col = cv[:, k]
        for a in range(0, n, 5):
            lines.append("".join(fortran_e23_16(x) for x in col[a:a + 5]))
    lines.append("# End")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def dv_sections(dv):
    """The four dv sections the reference writes for a perfect gas: p, T, c
    and dv(4, j) read from the 3-row array's storage (the next cell's p;
    0 past the last cell)."""
    dv = np.asarray(dv, np.float64)
    n = dv.shape[0]
    flat = np.ascontiguousarray(dv).reshape(-1)  # (p, T, c) per cell: the Fortran column-major dv(3, n)
    fourth = np.zeros(n)
    fourth[:n - 1] = flat[3:3 * n:3]
    return np.column_stack([dv[:, 0], dv[:, 1], dv[:, 2], fourth])


def write_dv_ascii(path, dv, time, res_init=0.0):
    """RFLU_WriteFlowASCII's .mixt.dva file for dv (nCells, 3) = (p, T, c)."""
    write_flow_ascii(path, dv_sections(dv), time, res_init)


def read_flow_ascii(path):
    """RFLU_ReadFlowASCII: returns dict(time, res_init, cv (nCells, nCv))."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    it = iter(lines)
    out = {}
    head = next(it)
    if not head.startswith("# ROCFLU flow file"):
        raise ValueError(f"{path}: not a Rocflu flow file")
    cv = None
    for ln in it:
        s = ln.strip()
        if s == "# Precision and range":
            prec, rng = (int(t) for t in next(it).split())
            if prec < PRECISION or rng < RANGE:
                raise ValueError(f"{path}: precision/range {prec}/{rng} below real*8")
        elif s == "# Initial residual":
            out["res_init"] = _parse_e(next(it))
        elif s == "# Physical time":
            out["time"] = _parse_e(next(it))
        elif s == "# Dimensions":
            n, nv = (int(t) for t in next(it).split())
            cv = np.zeros((n, nv))
        elif s == "# End":
            break
        elif s.startswith("#"):
            k = int(s[1:]) - 1
            vals = []
            while len(vals) < n:
                row = next(it)
                vals.extend(_parse_e(row[a:a + 23]) for a in range(0, len(row), 23) if row[a:a + 23].strip())
            cv[:, k] = vals[:n]
    out["cv"] = cv
    return out


def _rec(f, payload):
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


def _str(s):
    return s.ljust(CHRLEN).encode("ascii")


def write_flow_binary(path, cv, time, res_init=0.0):
    cv = np.asarray(cv, np.float64)
    n, nv = cv.shape
    with open(path, "wb") as f:
        _rec(f, _str("# ROCFLU flow file (v2)"))
        _rec(f, _str("# Precision and range"))
        _rec(f, struct.pack("<ii", PRECISION, RANGE))
        _rec(f, _str("# Initial residual"))
        _rec(f, struct.pack("<d", res_init))
        _rec(f, _str("# Physical time"))
        _rec(f, struct.pack("<d", time))
        _rec(f, _str("# Dimensions"))
        _rec(f, struct.pack("<ii", n, nv))
        for k in range(nv):
            _rec(f, _str(f"# {k + 1:2d}"))
            _rec(f, np.ascontiguousarray(cv[:, k], "<f8").tobytes())
        _rec(f, _str("# End"))


def read_flow_binary(path):
    """RFLU_ReadFlowBinary: returns dict(time, res_init, cv (nCells, nCv))."""
    data = open(path, "rb").read()
    pos, recs = 0, []
    while pos < len(data):
        (ln,) = struct.unpack_from("<i", data, pos)
        recs.append(data[pos + 4:pos + 4 + ln])
        (ln2,) = struct.unpack_from("<i", data, pos + 4 + ln)
        if ln2 != ln:
            raise ValueError(f"{path}: broken record markers at byte {pos}")
        pos += 8 + ln
    it = iter(recs)
    if not next(it).decode("ascii").startswith("# ROCFLU flow file"):
        raise ValueError(f"{path}: not a Rocflu flow file")
    out, cv, n = {}, None, 0
    for r in it:
        s = r.decode("ascii", errors="replace").strip()
        if s == "# Precision and range":
            next(it)
        elif s == "# Initial residual":
            (out["res_init"],) = struct.unpack("<d", next(it))
        elif s == "# Physical time":
            (out["time"],) = struct.unpack("<d", next(it))
        elif s == "# Dimensions":
            n, nv = struct.unpack("<ii", next(it))
            cv = np.zeros((n, nv))
        elif s == "# End":
            break
        elif s.startswith("#"):
            k = int(s[1:]) - 1
            cv[:, k] = np.frombuffer(next(it), "<f8", count=n)
    out["cv"] = cv
    return out


def fortran_list_real8(x):
    """gfortran list-directed REAL(8): F form with 17 significant digits for
    0.1 <= |x| < 1e16, else E form with a three-digit exponent."""
    x = float(x)
    ax = abs(x)
    if ax == 0.0:
        return "   0.0000000000000000     "
    if 0.1 <= ax < 1e16:
        ip = len(str(int(ax))) if ax >= 1 else 0
        return f"   {x:.{max(17 - ip, 0)}f}"
    s = f"{x:.16e}"
    mant, exp = s.split("e")
    e = int(exp)
    return f"   {mant}E{'+' if e >= 0 else '-'}{abs(e):03d}"


def append_restart_info(case, time, directory=""):
    """RFLU_WriteRestartInfo: append the physical time to <case>.rin."""
    with open(os.path.join(directory, f"{case}.rin"), "a") as f:
        f.write(fortran_list_real8(time) + "\n")


def read_restart_info(case, directory=""):
    """Times listed in <case>.rin (the last one is where a run resumes)."""
    with open(os.path.join(directory, f"{case}.rin")) as f:
        return [float(t) for t in f.read().split()]


# ---------------------------------------------------------------------------
# ppiclF particle files (libpicl/ppiclF/source/ppiclf_io.f:1362-1758)
# ---------------------------------------------------------------------------
LRS, LRP = 12, 48  # PPICLF_LRS, PPICLF_LRP (libpicl/user_files/PPICLF_USER.h)


def fortran_e14_7(x):
    """Fortran E14.7."""
    x = float(x)
    if x == 0.0:
        return " 0.0000000E+00"
    m = f"{x:.6e}"
    mant, exp = m.split("e")
    sign = "-" if mant.startswith("-") else ""
    digits = mant.lstrip("-").replace(".", "")
    e = int(exp) + 1
    return f"{sign}0.{digits}E{'+' if e >= 0 else '-'}{abs(e):02d}".rjust(14)


def write_particle_vtu(path, y, rprop=None, tags=None, time=0.0, cycle=0):
    """ppiclf_io_WriteParticleVTU: an UnstructuredGrid of points with
    Float32 arrays Position, y01-y12, rprop01-rprop48 and tag01-tag03 in
    raw appended data (each block a 4-byte byte count, then the values).
    y: (n, 12); rprop: (n, 48) or None (zeros); tags: (n, 3) or None."""
    y = np.asarray(y, np.float64)
    n = y.shape[0]
    rprop = np.zeros((n, LRP)) if rprop is None else np.asarray(rprop, np.float64)
    tags = np.zeros((n, 3)) if tags is None else np.asarray(tags, np.float64)
    arrays = [("Position", 3, y[:, :3])]
    arrays += [(f"y{j + 1:02d}", 1, y[:, j]) for j in range(LRS)]
    arrays += [(f"rprop{j + 1:02d}", 1, rprop[:, j]) for j in range(LRP)]
    arrays += [(f"tag{j + 1:02d}", 1, tags[:, j]) for j in range(3)]
    head = ['<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian">',
            ' <UnstructuredGrid>', '  <FieldData>',
            '   <DataArray type="Float32" Name="TIME" NumberOfTuples="1" format="ascii"> '
            f'{fortran_e14_7(time)} </DataArray> ',
            '   <DataArray type="Int32" Name="CYCLE" NumberOfTuples="1" format="ascii"> '
            f'{int(cycle)} </DataArray> ',
            '  </FieldData>', f'  <Piece NumberOfPoints="{n}" NumberOfCells="0"> ']
    off = 0

    def da(name, nc):
        return (f'    <DataArray type="Float32" Name="{name}" NumberOfComponents="{nc}" '
                f'format="append" offset="{off}"/>')
    head.append('   <Points>')
    head.append(da("Position", 3))
    off += 3 * 4 * n + 4
    head.append('   </Points>')
    head.append('   <PointData>')
    for name, nc, _ in arrays[1:]:
        head.append(da(name, nc))
        off += nc * 4 * n + 4
    head += ['   </PointData> ', '   <Cells> ',
             '    <DataArray type="Int32" Name="connectivity" format="ascii"/> ',
             '    <DataArray type="Int32" Name="offsets" format="ascii"/> ',
             '    <DataArray type="Int32" Name="types" format="ascii"/> ',
             '   </Cells> ', '  </Piece> ', ' </UnstructuredGrid> ']
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode("ascii"))
        f.write(b' <AppendedData encoding="raw">_')
        for name, nc, a in arrays:
            blob = np.ascontiguousarray(a, "<f4").tobytes()
            f.write(struct.pack("<i", len(blob)))
            f.write(blob)
        f.write(b"</AppendedData>\n</VTKFile>\n")


def read_particle_vtu(path):
    """ppiclf_io_ReadParticleVTU's counterpart: returns dict(time, cycle,
    y (n, 12), rprop (n, 48), tags (n, 3)) as float64 (the file is Float32)."""
    import re
    data = open(path, "rb").read()
    cut = data.index(b'<AppendedData encoding="raw">_') + len(b'<AppendedData encoding="raw">_')
    head = data[:cut].decode("ascii")
    n = int(re.search(r'NumberOfPoints="(\d+)"', head).group(1))
    time = float(re.search(r'Name="TIME"[^>]*>\s*([^<\s]+)', head).group(1))
    cycle = int(re.search(r'Name="CYCLE"[^>]*>\s*([^<\s]+)', head).group(1))
    arrs = {}
    for name, nc, off in re.findall(r'Name="([^"]+)" NumberOfComponents="(\d+)" format="append" offset="(\d+)"', head):
        p = cut + int(off)
        (nb,) = struct.unpack_from("<i", data, p)
        arrs[name] = np.frombuffer(data, "<f4", count=nb // 4, offset=p + 4).astype(np.float64).reshape(n, int(nc))
    y = np.concatenate([arrs[f"y{j + 1:02d}"] for j in range(LRS)], axis=1)
    rprop = np.concatenate([arrs[f"rprop{j + 1:02d}"] for j in range(LRP)], axis=1)
    tags = np.concatenate([arrs[f"tag{j + 1:02d}"] for j in range(3)], axis=1)
    return dict(time=time, cycle=cycle, y=y, rprop=rprop, tags=tags, position=arrs["Position"])
