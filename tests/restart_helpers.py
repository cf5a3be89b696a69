"""Test-side writer of SuLi restart records (the list restart_salva writes,
SuLi-IC/11_diversos.f90:593-595, shapes from SuLi-IC/0_modules.f90) in the
list-directed styles compilers produce: gfortran's fixed 17-digit fields
with r*c repeat counts, D exponents, comma separators, bare signed
exponents, null values.  Used to check suli_ps_restart_read."""
import numpy as np


def record_shapes(nx, ny, nz, byxf_full=True):
    nbox = (nx + 2, ny + 2, nz + 2)
    n = (nx, ny, nz)
    yz, y1z, yz1 = (ny + 2, nz + 2), (ny + 3, nz + 2), (ny + 2, nz + 3)
    xz, x1z, xz1 = (nx + 2, nz + 2), (nx + 3, nz + 2), (nx + 2, nz + 3)
    xy, x1y, xy1 = (nx + 2, ny + 2), (nx + 3, ny + 2), (nx + 2, ny + 3)
    return [
        ("u", (nx + 3, ny + 2, nz + 2)), ("v", (nx + 2, ny + 3, nz + 2)), ("w", (nx + 2, ny + 2, nz + 3)),
        ("ku", x1y), ("kv", xy1), ("kw", xy), ("prd0", nbox), ("prd1", nbox), ("prd", nbox),
        ("rho", n), ("ls_mu", n), ("ls", n),
        ("bxx0", yz), ("bxx1", yz), ("bxy0", y1z), ("bxz0", yz1), ("bxxf", yz), ("bxxf1", yz),
        ("bxyf", y1z), ("bxzf", yz1), ("byx0", x1z), ("byy0", xz), ("byy1", xz), ("byz0", xz1),
        ("byxf", x1z if byxf_full else (0,)), ("byyf", xz), ("byyf1", xz), ("byzf", xz1),
        ("bzx0", x1y), ("bzy0", xy1), ("bzz0", xy), ("bzz1", xy), ("bzxf", x1y), ("bzyf", xy1),
        ("bzzf", xy), ("bzzf1", xy), ("ub", (nx + 3, ny + 2, nz + 2)), ("vb", (nx + 2, ny + 3, nz + 2)),
        ("wb", (nx + 2, ny + 2, nz + 3)), ("d_max", (1,)), ("d_min", (1,)), ("ls_m", (1,)),
        ("hsx", n), ("hsy", n), ("hsz", n), ("hs", n), ("kurv", n),
        ("xmu", (nx + 1, ny, nz)), ("ymu", (nx, ny + 1, nz)), ("zmu", (nx, ny, nz + 1)),
        ("tf_u", (nx + 1, ny, nz)), ("tf_v", (nx, ny + 1, nz)), ("tf_w", (nx, ny, nz + 1)),
        ("vol_ini", (1,)), ("vol_ins", (1,)),
    ]


def _fmt(x, style, k):
    if style == "gfortran":
        s = f"{x:.16E}"                       # 1.2345678901234567E-02
        mant, ex = s.split("E")
        e = int(ex)
        if abs(e) >= 100:                     # Fortran drops the letter for 3-digit exponents
            return f"{mant}{e:+d}"
        s = f"{mant}E{e:+04d}"
        if k % 5 == 3:
            s = s.replace("E", "D")
        return s
    return repr(float(x))


def write_record(path, arrays, nx, ny, nz, byxf_full=True, style="gfortran", nulls=0, slash=False):
    """arrays: name -> Fortran-ordered array (missing ones are zeros).  The
    first `nulls` values of u are written as a null repeat (`r*`)."""
    vals = []
    for name, shape in record_shapes(nx, ny, nz, byxf_full):
        a = arrays.get(name)
        if a is None:
            a = np.zeros(shape)
        vals.append(np.asarray(a, dtype=np.float64).ravel(order="F"))
    flat = np.concatenate(vals)
    out = []
    i = 0
    if nulls:
        out.append(f"{nulls}*")
        i = nulls
    k = 0
    while i < flat.size:
        j = i
        while j + 1 < flat.size and flat[j + 1] == flat[i]:
            j += 1
        run = j - i + 1
        body = _fmt(flat[i], style, k)
        out.append(f"{run}*{body}" if run > 1 else body)
        i = j + 1
        k += 1
    with open(path, "w") as f:
        line = []
        for t, tok in enumerate(out):
            line.append(tok)
            if len(line) == 4:
                sep = "," if t % 3 == 0 else "   "
                f.write(" " + sep.join(line) + "\n")
                line = []
        if line:
            f.write(" " + "   ".join(line) + "\n")
        if slash:
            f.write(" /\n garbage after the record\n")
    return flat.size
