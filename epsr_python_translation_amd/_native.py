"""Build and load libepsr_mi355x.so (the HIP kernels + C ABI) for gfx950.

The library is built in-tree next to this file.  Loading never falls back to anything:
if the shared library is missing or fails to load, the GPU entry points raise.
"""

from __future__ import annotations

import ctypes as C
import os
import pathlib
import shutil
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB_PATH = HERE / "libepsr_mi355x.so"
INCLUDE = HERE.parent / "include"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # the reference's float32 contract: no contraction to FMA, correctly rounded
    # division and square root, no fast-math
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = False, jobs: int | None = None) -> pathlib.Path:
    """Compile every source to an object (in parallel, each its own translation unit as in
    a one-command build) and link the shared library."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    tmp = LIB_PATH.with_suffix(".so.tmp")
    objdir = HERE / "build"
    objdir.mkdir(exist_ok=True)
    env = dict(os.environ)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-6000:]}")

    srcs = sources()
    objs = [objdir / (p.stem + ".o") for p in srcs]
    n = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(n) as ex:
        list(ex.map(run, [[hipcc, *HIPCC_FLAGS, f"-I{INCLUDE}", "-c", "-o", str(o), str(p)]
                          for p, o in zip(srcs, objs)]))
    run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


_LIB = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or _native.build())")
        _LIB = C.CDLL(str(LIB_PATH))
        _declare(_LIB)
    return _LIB


vp = C.c_void_p
ip = C.POINTER(C.c_int)
i32p = C.POINTER(C.c_int32)
fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)


class BoxDesc(C.Structure):
    _fields_ = [("nmol", C.c_int), ("natom", C.c_int), ("ntype", C.c_int), ("npairs", C.c_int),
                ("nmoltype", C.c_int), ("itype", ip), ("naf", ip), ("nal", ip), ("mtype", ip),
                ("npairref", ip), ("n1extr", ip), ("n2extr", ip), ("cfrac", fp),
                ("natomr", C.c_float), ("sizefactor", C.c_float), ("dmax", C.c_float),
                ("cellst", C.c_float)]


class GensqBox(C.Structure):
    _fields_ = [("npairs", C.c_int), ("ldw", C.c_int), ("ldq", C.c_int), ("wav", vp), ("natomsphere", vp),
                ("sofq", vp)]


class StogBox(C.Structure):
    _fields_ = [("sofq", vp), ("gofr", vp), ("nqmin", vp), ("nqmax", vp), ("xwt", vp), ("ncols", C.c_int),
                ("ldg", C.c_int)]


class Coords(C.Structure):
    _fields_ = [("xa", vp), ("ya", vp), ("za", vp), ("xadd", vp), ("yadd", vp), ("zadd", vp)]


class Tables(C.Structure):
    _fields_ = [("ld", C.c_int), ("npt2", C.c_int), ("xstar", vp), ("xavs", vp), ("xavs1", vp), ("ecoul", vp),
                ("ecoul1", vp)]


class MCParams(C.Structure):
    _fields_ = [("simt", C.c_float), ("stepr", C.c_float), ("stepm", C.c_float), ("rejrate", C.c_float),
                ("ireset", C.c_int)]


class MCStats(C.Structure):
    _fields_ = [("moves", C.c_int), ("accepted", C.c_int), ("exact_resolutions", C.c_int), ("retries", C.c_int),
                ("ms_pairs", C.c_double), ("ms_resolve", C.c_double), ("pushed_entries", C.c_longlong),
                ("resolver_bytes", C.c_double), ("stage1_rejected", C.c_int), ("resolver_passes", C.c_int),
                ("cycles", C.c_int), ("accepted_total", C.c_longlong)]


LAUNCH_ATOGEN_WORKGROUPS = 1
LAUNCH_PAIR_BLOCK = 2
LAUNCH_RESOLVE_PASSES = 3


class DescParams(C.Structure):
    _fields_ = [("ng", C.c_int), ("rc", C.c_float), ("r0", C.c_float), ("dmu", C.c_float), ("width", C.c_float),
                ("xscale", C.c_float)]


class MLP(C.Structure):
    _fields_ = [("d", C.c_int), ("h", C.c_int), ("w", vp), ("wb", vp), ("m", vp), ("v", vp), ("step", C.c_int)]


class DA(C.Structure):
    _fields_ = [("mlp", C.POINTER(MLP)), ("desc", DescParams), ("seed", C.c_ulonglong), ("des", vp)]


class MCRecord(C.Structure):
    _fields_ = [("desc", C.POINTER(DescParams)), ("ldx", C.c_int), ("x", vp), ("target", vp), ("des", vp),
                ("pass_", vp)]


class IntraDesc(C.Structure):
    _fields_ = [("ldb", C.c_int), ("qtype", ip), ("imageatom", ip), ("nbond", ip), ("iabond", ip), ("tbond", fp),
                ("amassr", fp), ("amassa", fp), ("ngroup", C.c_int), ("nlabrot", C.c_int), ("ldd", C.c_int),
                ("nrf", ip), ("nrl", ip), ("labaxi", ip), ("rotf", ip), ("rotl", ip), ("labrot", ip), ("isdih", ip),
                ("ndihedralang", ip), ("labrotdih", ip), ("dihedralang", fp), ("dmaxmol", C.c_float),
                ("tempb", C.c_float), ("refine_intra_fac", C.c_float), ("refine_intra", C.c_int)]


class MCFullParams(C.Structure):
    _fields_ = [("simt", C.c_float), ("stepr", C.c_float), ("stepm", C.c_float), ("stepri", C.c_float),
                ("stepmi", C.c_float), ("rejrate", C.c_float), ("tempangwt", C.c_float), ("tempdihwt", C.c_float),
                ("ireset", C.c_int), ("intra", C.c_int), ("rotfreq", C.c_int)]


class DMSDesc(C.Structure):
    """epsr_dms_desc_t"""
    _fields_ = [("nfile", C.c_int), ("nq", C.c_int), ("ldq", C.c_int), ("npairs", C.c_int), ("nxfile", C.c_int),
                ("q", C.c_void_p), ("datai", C.c_void_p), ("soqintra", C.c_void_p), ("soqinter", C.c_void_p),
                ("weightx", C.c_void_p), ("x", C.c_void_p), ("ciwt", C.c_void_p), ("cwt", C.c_void_p),
                ("xfile", C.c_void_p), ("nqmind", C.c_void_p), ("nqmaxd", C.c_void_p), ("nqmin", C.c_void_p),
                ("nqmax", C.c_void_p), ("szeros", C.c_void_p), ("porodamp", C.c_void_p), ("rshmin", C.c_void_p),
                ("porodq", C.c_float), ("porodpwr", C.c_float), ("nsmoop", C.c_int), ("npt1", C.c_int),
                ("rho", C.c_float)]


class MCSteps(C.Structure):
    _fields_ = [("wholerot", fp), ("wholetrans", fp), ("internrot", fp), ("interntrans", fp), ("nwhole", ip),
                ("nmovem", ip), ("ninter", ip), ("nmovei", ip)]


# name -> (restype, argtypes); every symbol declared in include/epsr_mi355x.h
SIGNATURES = {
    "epsr_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "epsr_ctx_destroy": (None, [vp]),
    "epsr_last_error": (C.c_char_p, [vp]),
    "epsr_ctx_set_stream": (C.c_int, [vp, vp]),
    "epsr_ctx_synchronize": (C.c_int, [vp]),
    "epsr_ctx_last_kernel_ms": (C.c_double, [vp, C.c_char_p]),
    "epsr_ctx_kernel_ms_sum": (C.c_double, [vp, C.c_char_p, C.POINTER(C.c_longlong)]),
    "epsr_ctx_enable_timing": (C.c_int, [vp, C.c_int]),
    "epsr_box_create": (C.c_int, [vp, C.POINTER(BoxDesc), C.POINTER(vp)]),
    "epsr_box_destroy": (None, [vp]),
    "epsr_box_set_launch": (C.c_int, [vp, C.c_int, C.c_int]),
    "epsr_graph_begin": (C.c_int, [vp]),
    "epsr_graph_end": (C.c_int, [vp, C.POINTER(vp)]),
    "epsr_graph_launch": (C.c_int, [vp, vp]),
    "epsr_graph_destroy": (None, [vp]),
    "epsr_ran1_stream": (C.c_int, [i32p, C.c_int, fp]),
    "epsr_atogen": (C.c_int, [vp, vp, C.POINTER(Coords), C.c_int, C.c_int, vp, vp, vp, ip, i32p, vp, vp, vp, vp]),
    "epsr_atogen_bin_check": (C.c_int, [vp, C.c_float, C.c_int, C.c_float, C.c_float, C.c_uint32, C.c_uint32,
                                        vp]),
    "epsr_gensq": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int, vp, vp, C.c_int, C.c_int]),
    "epsr_gensq_noalloc": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int, vp, vp,
                                     C.c_int]),
    "epsr_stog": (C.c_int, [vp, C.c_float, vp, vp, C.c_int, vp, vp, C.c_int, vp, vp, C.c_int, C.c_int, vp,
                            C.c_int]),
    "epsr_gensq_batch": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, vp]),
    "epsr_stog_batch": (C.c_int, [vp, C.c_float, vp, C.c_int, vp, C.c_int, C.c_int, vp]),
    "epsr_ledger_create": (C.c_int, [vp, vp, C.POINTER(vp)]),
    "epsr_ledger_destroy": (None, [vp]),
    "epsr_cells_init": (C.c_int, [vp, vp, vp, C.POINTER(Coords), C.c_int, C.c_float]),
    "epsr_fmolec1": (C.c_int, [vp, vp, C.POINTER(Coords), C.POINTER(Tables), vp, vp, dp, dp, fp, fp, ip, ip]),
    "epsr_ledger_export": (C.c_int, [vp, vp, C.c_int, ip, ip, ip, fp, ip]),
    "epsr_en_change": (C.c_int, [vp, vp, C.POINTER(Coords), C.POINTER(Tables), vp, C.c_int, vp, fp]),
    "epsr_en_mole_pair": (C.c_int, [vp, vp, C.POINTER(Coords), C.POINTER(Tables), C.c_int, vp, vp, vp]),
    "epsr_sumsq": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, C.c_double, C.c_double, vp, vp]),
    "epsr_sumgr": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, C.c_double, C.c_double, vp, vp, vp]),
    "epsr_invert_cwt": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, C.c_float, vp, C.c_int, C.c_int,
                                  vp, ip]),
    "epsr_press": (C.c_int, [vp, vp, C.POINTER(Coords), C.POINTER(Tables), vp, C.c_int, C.c_float, C.c_float, fp, fp,
                             C.POINTER(C.c_double), C.c_int, fp, fp]),
    "epsr_molfit": (C.c_int, [vp, vp, C.POINTER(Coords), C.POINTER(Tables), vp, C.POINTER(MCParams), fp, fp, ip, ip,
                              vp, i32p, fp, C.POINTER(MCStats)]),
    "epsr_molfit_cycles": (C.c_int, [vp, vp, C.POINTER(Coords), C.POINTER(Tables), vp, C.POINTER(MCParams), C.c_int,
                                     fp, fp, ip, ip, vp, i32p, fp, C.POINTER(MCStats)]),
    "epsr_molfit_da": (C.c_int, [vp, vp, C.POINTER(Coords), C.POINTER(Tables), vp, C.POINTER(MCParams), fp, fp, ip,
                                 ip, vp, i32p, fp, C.POINTER(DA), C.POINTER(MCRecord), C.POINTER(MCStats)]),
    "epsr_box_set_intra": (C.c_int, [vp, C.POINTER(IntraDesc)]),
    "epsr_en_mole_intra": (C.c_int, [vp, vp, C.POINTER(Coords), C.POINTER(Tables), vp, vp, C.c_float, C.c_float,
                                     C.c_float, vp]),
    "epsr_molfit_full": (C.c_int, [vp, vp, C.POINTER(Coords), C.POINTER(Tables), vp, vp, vp, C.POINTER(MCFullParams),
                                   C.POINTER(MCSteps), vp, i32p, fp, fp, ip, C.POINTER(MCStats)]),
    "epsr_descriptors": (C.c_int, [vp, vp, C.POINTER(Coords), C.POINTER(Coords), vp, C.POINTER(DescParams), C.c_int,
                                   vp, vp]),
    "epsr_mlp_nparams": (C.c_longlong, [C.c_int, C.c_int]),
    "epsr_mlp_sync": (C.c_int, [vp, C.POINTER(MLP)]),
    "epsr_mlp_forward": (C.c_int, [vp, C.POINTER(MLP), C.c_int, vp, C.c_int, vp]),
    "epsr_mlp_grad": (C.c_int, [vp, C.POINTER(MLP), C.c_int, vp, C.c_int, vp, C.c_float, vp, fp]),
    "epsr_mlp_adam": (C.c_int, [vp, C.POINTER(MLP), vp, C.c_float, C.c_float, C.c_float, C.c_float]),
    "epsr_truncation": (C.c_int, [vp, C.c_int, vp, C.c_float, C.c_float, vp, vp]),
    "epsr_generate_rp": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int]),
    "epsr_generate_ep": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, C.c_int, ip, C.c_int, fp, C.c_int, vp, vp, vp, vp,
                                   vp, vp, C.c_int, fp, fp]),
    "epsr_calc_pwrexpq": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, vp,
                                    vp, C.c_float, ip]),
    "epsr_foqols": (C.c_int, [vp, C.c_int, vp, C.c_int, vp, vp, vp, C.c_float, C.c_float, C.c_float, C.c_float,
                              C.c_float, C.c_float, vp, C.c_int]),
    "epsr_expand_all": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, C.c_float, vp]),
    "epsr_smoop": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp]),
    "epsr_fit_expansion": (C.c_int, [vp, C.c_int, vp, C.c_int, vp, C.c_longlong, vp, C.c_int, vp, vp, C.c_int, fp,
                                     fp, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, i32p, C.c_int,
                                     ip]),
    "epsr_ptgnsq_diff": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, vp, fp, fp, ip, ip, C.c_float,
                                   vp, vp, ip, C.c_int]),
    "epsr_data_minus_sim": (C.c_int, [vp, C.POINTER(DMSDesc), vp, vp, vp, vp]),
    "epsr_ptgnsq_rfactor": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, fp]),
    "epsr_ptgnsq_update": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, fp, ip, fp, fp, C.c_int, C.c_float, fp, fp,
                                     C.c_int]),
}

FIT_MC = 0
FIT_NORMAL = 1


def _declare(L):
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def header_symbols():
    """Function names declared in include/*.h (for the export test)."""
    import re

    names = set()
    for h in INCLUDE.glob("*.h"):
        text = h.read_text()
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(epsr_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names
