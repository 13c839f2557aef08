"""ctypes view of the reference-compatible C interface (include/FFTSVDpbeAPI.h)
used by the tests: struct layouts, globals and the solve-path entry points."""
import ctypes as C

import numpy as np

import fftsvd_amd as fa


class PDBentry(C.Structure):
    """FFTSVDpbeAPI.h:23-42 (reference), include/FFTSVDpbeAPI.h"""
    _fields_ = [("record", C.c_char * 7), ("atomnumber", C.c_uint), ("atomname", C.c_char * 5),
                ("alternatelocation", C.c_char), ("residuename", C.c_char * 4), ("chain", C.c_char),
                ("residuenumber", C.c_uint), ("residueinsertion", C.c_char), ("x", C.c_double),
                ("y", C.c_double), ("z", C.c_double), ("occupancy", C.c_double), ("temperature", C.c_double),
                ("footnotenumber", C.c_uint), ("radius", C.c_double), ("charge", C.c_double),
                ("potential", C.c_double)]


class SIZentry(C.Structure):
    _fields_ = [("atomlabel", C.c_char * 7), ("residuelabel", C.c_char * 4), ("radius", C.c_double)]


class CRGentry(C.Structure):
    _fields_ = [("atomlabel", C.c_char * 7), ("residuelabel", C.c_char * 4), ("residuenumber", C.c_uint),
                ("chain", C.c_char), ("charge", C.c_double)]


class Charge(C.Structure):
    _fields_ = [("numcharges", C.c_uint), ("globalindexstart", C.c_uint), ("points", C.c_void_p),
                ("charges", C.POINTER(C.c_double))]


class SurfaceOperator(C.Structure):
    pass


SurfaceOperator._fields_ = [("kernel", C.c_int), ("epsilon", C.c_double), ("kappa", C.c_double),
                            ("mynumpanels", C.c_uint), ("startphi", C.c_uint), ("startdphi", C.c_uint),
                            ("tree", C.c_void_p), ("M3", C.c_void_p), ("numchildren", C.c_uint),
                            ("children", C.POINTER(C.POINTER(SurfaceOperator))),
                            ("resultInternal", C.POINTER(C.c_double)), ("resultExternal", C.POINTER(C.c_double)),
                            ("parent", C.POINTER(SurfaceOperator)), ("charges", C.POINTER(Charge))]


class Surfaces:
    """Outputs of readSRF."""

    def __init__(self):
        self.ptrs = [C.c_void_p() for _ in range(4)]            # Panel** per surface class
        self.counts = [C.POINTER(C.c_uint)() for _ in range(4)]  # panels per surface
        self.num = [C.c_uint() for _ in range(4)]                # surfaces per class
        self.parents = [C.POINTER(C.c_uint)() for _ in range(3)]
        self.total = C.c_uint()


def lib():
    return fa.lib()


def set_global(name, value, ctype=C.c_double):
    ctype.in_dll(lib(), name).value = value


def get_global(name, ctype=C.c_double):
    return ctype.in_dll(lib(), name).value


def read_srf(path):
    L = lib()
    L.readSRF.restype = None
    s = Surfaces()
    b = C.byref
    L.readSRF(str(path).encode(), b(s.ptrs[0]), b(s.counts[0]), b(s.num[0]), b(s.ptrs[1]), b(s.counts[1]),
              b(s.num[1]), b(s.parents[0]), b(s.ptrs[2]), b(s.counts[2]), b(s.num[2]), b(s.parents[1]),
              b(s.ptrs[3]), b(s.counts[3]), b(s.num[3]), b(s.parents[2]), b(s.total))
    return s


def solve_path(s, pdb):
    """generateSurfaceOperator -> preconditioner -> RHS -> solve ->
    collectPotentials, exactly as the reference driver calls them
    (FFTSVDpbe.c:179-234).  Returns the operator (ctypes pointer)."""
    L = lib()
    for f in ("generateSurfaceOperator", "generateSurfaceOperatorPreconditioner", "generateRHS",
              "solveSurfaceOperator", "collectPotentials"):
        getattr(L, f).restype = None
    op = C.POINTER(SurfaceOperator)()
    n = len(pdb)
    L.generateSurfaceOperator(C.byref(op), pdb, C.c_uint(n), s.ptrs[0], s.counts[0], s.num[0], s.ptrs[1],
                              s.counts[1], s.num[1], s.parents[0], s.ptrs[2], s.counts[2], s.num[2], s.parents[1],
                              s.ptrs[3], s.counts[3], s.num[3], s.parents[2], s.total)
    pc = C.c_void_p()
    L.generateSurfaceOperatorPreconditioner(C.byref(pc), op, s.total)
    rhs = C.POINTER(C.c_double)()
    L.generateRHS(C.byref(rhs), op, s.total)
    sol = C.POINTER(C.c_double)()
    L.solveSurfaceOperator(C.byref(sol), op, pc, rhs, s.total)
    L.collectPotentials(pdb, C.c_uint(n), pdb, C.c_uint(n), op, sol)
    return op, rhs, sol


def qual_path(s, pdb):
    """generateQualocationOperator -> preconditioner -> RHS -> solve ->
    collectQualocationPotentials, as the reference driver calls them with
    usequalocation (FFTSVDpbe.c:169-231).  Returns (rhs, sol) pointers."""
    L = lib()
    for f in ("generateQualocationOperator", "generateQualocationOperatorPreconditioner", "generateQualocationRHS",
              "solveQualocationOperator", "collectQualocationPotentials"):
        getattr(L, f).restype = None
    qo = C.c_void_p()
    n = len(pdb)
    L.generateQualocationOperator(C.byref(qo), pdb, C.c_uint(n), s.ptrs[1], s.counts[1], s.num[1], s.ptrs[2],
                                  s.counts[2], s.num[2], s.parents[1], s.total)
    pc = C.c_void_p()
    L.generateQualocationOperatorPreconditioner(C.byref(pc), qo, s.total)
    rhs = C.POINTER(C.c_double)()
    L.generateQualocationRHS(C.byref(rhs), qo, s.total)
    sol = C.POINTER(C.c_double)()
    L.solveQualocationOperator(C.byref(sol), qo, pc, rhs, s.total)
    L.collectQualocationPotentials(pdb, C.c_uint(n), pdb, C.c_uint(n), qo, sol)
    return rhs, sol


def solve_bem_problem(verts9, eps_in, eps_out, rhs):
    """solveBEMproblem (FFTSVDpbeAPI.c:1974-2026), the GAMESS entry."""
    L = lib()
    L.solveBEMproblem.restype = None
    v = np.ascontiguousarray(verts9, dtype=np.float64).reshape(-1)
    r = np.ascontiguousarray(rhs, dtype=np.float64)
    sol = np.zeros(len(r))
    n = C.c_uint(len(r))
    ei, eo = C.c_double(eps_in), C.c_double(eps_out)
    dp = C.POINTER(C.c_double)
    L.solveBEMproblem(C.byref(n), v.ctypes.data_as(dp), C.byref(ei), C.byref(eo), r.ctypes.data_as(dp),
                      sol.ctypes.data_as(dp))
    return sol


def free(p):
    C.CDLL(None).free(C.cast(p, C.c_void_p))


def write_srf(path, dielectric_names, salt_names=(), parents=None, cavity_names=(), cavity_parents=(),
              salt_cavity_names=(), salt_cavity_parents=()):
    """.srf with flat surfaces: salt, dielectric, dielectric cavity and
    salt cavity lines, then the dielectric, cavity and salt cavity parents
    (FFTSVDpbeAPI.c:776-1032 grammar)."""
    parents = parents if parents is not None else [0] * len(dielectric_names)
    with open(path, "w") as f:
        f.write("f\nf\n")
        f.write(" ".join(salt_names) + "\n")
        f.write(" ".join(dielectric_names) + "\n")
        f.write(" ".join(cavity_names) + "\n")
        f.write(" ".join(salt_cavity_names) + "\n")
        f.write(" ".join(str(p) for p in parents) + "\n")
        f.write(" ".join(str(p) for p in cavity_parents) + "\n")
        f.write(" ".join(str(p) for p in salt_cavity_parents) + "\n")
