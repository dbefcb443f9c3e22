#!/usr/bin/env python3
"""Build the reference's own translation units into oracle/_ref/ (test infrastructure).

The reference selects Scheme / Riemann / Limiter / Boundary at link time
(Test/Sod/Makefile:1-45). This recipe compiles, for each combination the tests
use, the reference sources where they lie under /root/reference together with
oracle/ref_harness.cpp, oracle/ref_limiter.cpp and the chosen problem hooks, into
oracle/_ref/libref_<case>_<scheme>_<riemann>_<limiter>_<boundary>.so.
SimpleBound's sources still define the single-device signatures; for it
oracle/ref_simplebound.cpp adds the current signatures, forwarding to them.

Compiler: g++ -O2 -fopenacc (OpenACC host fallback, sequential) with
-ffp-contract=off, so the fixture arithmetic is plain IEEE double in source order.
Nothing is copied out of /root/reference; when it is absent this script exits
without building (a machine without the reference sources uses the libraries
built where they were present).

Usage: python3 oracle/build_ref.py [-j N] [--list]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")

SRC = REF + "/Src"
TEST = REF + "/Test"

CASE_HOOKS = {
    # case -> hook sources
    "sodx": [os.path.join(HERE, "ref_hooks_sodx.cpp"), TEST + "/Sod/Level_Set_function.cpp",
             TEST + "/Sod/Global_Boudary.cpp"],
    "sod": [TEST + "/Sod/Initialize.cpp", TEST + "/Sod/Level_Set_function.cpp", TEST + "/Sod/Global_Boudary.cpp"],
    "incline": [TEST + "/InclineSod/Initialize.cpp", TEST + "/InclineSod/Level_Set_function.cpp",
                TEST + "/InclineSod/Global_Boudary.cpp"],
    "vortex": [TEST + "/Vortex/Initialize.cpp", TEST + "/Vortex/Level_Set_function.cpp",
               TEST + "/Vortex/Global_Boudary.cpp"],
    # Test/PNGGeo's initialiser and edge hook (single-device signatures, adapted by
    # oracle/ref_pnggeo.cpp) over InclineSod's level set (its own mask level set is
    # built on its own: libref_pngmask.so)
    "pngshock": [TEST + "/PNGGeo/Initialize.cpp", TEST + "/PNGGeo/Global_Boudary.cpp",
                 TEST + "/InclineSod/Level_Set_function.cpp", os.path.join(HERE, "ref_pnggeo.cpp")],
}
NUM_SCH = {"Regular": 1, "IDWBaseGFM": 8, "LinearBaseGFM": 4, "RiemannBaseGFM": 4, "BCGFM": 16, "HBCGFM": 16,
           "SimpleBound": 2}

SCHEMES = ["MUSCL", "MHM", "Godunov", "Lax"]
RIEMANNS = ["HLL", "HLLC", "PVRS", "StegerWarming"]
LIMITERS = ["minmod", "VanLeer"]


def combos():
    out = []
    for s in SCHEMES:
        for r in ["HLL", "PVRS", "StegerWarming"]:
            for l in LIMITERS:
                if s == "Lax" and (r != "HLL" or l != "minmod"):
                    continue
                out.append(("sodx", s, r, l, "Regular"))
    out.append(("sodx", "MUSCL", "HLLC", "minmod", "Regular"))
    out.append(("sod", "MUSCL", "HLL", "minmod", "Regular"))
    for b in ["IDWBaseGFM", "LinearBaseGFM", "RiemannBaseGFM", "BCGFM", "HBCGFM"]:
        out.append(("incline", "MUSCL", "HLL", "minmod", b))
    out.append(("incline", "MHM", "HLL", "minmod", "IDWBaseGFM"))
    # PVRS in 2-D: the inclined channel, and the row-split hook with a
    # downward-moving upper state, whose collision takes the y solver's
    # branch with the stray store (PVRS/Riemann_solver.cpp:541-551)
    for s in ["MUSCL", "MHM"]:
        out.append(("incline", s, "PVRS", "minmod", "IDWBaseGFM"))
        out.append(("sod", s, "PVRS", "minmod", "Regular"))
    out.append(("incline", "MUSCL", "StegerWarming", "minmod", "IDWBaseGFM"))
    out.append(("vortex", "MUSCL", "HLL", "minmod", "IDWBaseGFM"))
    out.append(("incline", "MUSCL", "HLL", "minmod", "SimpleBound"))
    out.append(("pngshock", "MUSCL", "StegerWarming", "minmod", "IDWBaseGFM"))
    out.append(("pngshock", "MUSCL", "HLL", "minmod", "IDWBaseGFM"))
    out.append(("vortex", "MUSCL", "HLL", "minmod", "SimpleBound"))
    # Lax in 2-D: the row-split Sod hook and the inclined channel with IDW ghost fluid
    out.append(("sod", "Lax", "HLL", "minmod", "Regular"))
    out.append(("incline", "Lax", "HLL", "minmod", "IDWBaseGFM"))
    # more scheme / solver / limiter combinations on the IDW channel
    for s, r, l in [("Godunov", "HLL", "minmod"), ("Godunov", "StegerWarming", "minmod"),
                    ("MUSCL", "HLL", "VanLeer"), ("MHM", "StegerWarming", "minmod"), ("MHM", "HLL", "VanLeer")]:
        out.append(("incline", s, r, l, "IDWBaseGFM"))
    return out


def lib_name(c):
    return "libref_%s_%s_%s_%s_%s.so" % c


def sources(c):
    case, scheme, riemann, limiter, bound = c
    srcs = [SRC + "/Geometry/Psy_coord.cpp", SRC + "/Geometry/Level_Set.cpp", SRC + "/Geometry/Mirror_IDX.cpp",
            SRC + "/Utils/ParamReader.cpp", SRC + "/Utils/CaculateMass.cpp",
            SRC + "/Base/Common/ComputeDt.cpp", SRC + "/Base/Common/Conserve2Flux.cpp",
            SRC + "/Base/TimeAdvance/SingleStep/SingleStep.cpp"]
    bdir = SRC + "/BoundaryCondition/" + bound
    srcs += sorted(os.path.join(bdir, f) for f in os.listdir(bdir) if f.endswith(".cpp"))
    adir = SRC + "/Base/Euler/Advection/" + scheme
    srcs += sorted(os.path.join(adir, f) for f in os.listdir(adir) if f.endswith(".cpp"))
    if scheme != "Lax":
        rdir = SRC + "/Base/Euler/Riemann/" + riemann
        srcs += sorted(os.path.join(rdir, f) for f in os.listdir(rdir) if f.endswith(".cpp"))
    srcs += CASE_HOOKS[case]
    srcs += [os.path.join(HERE, "ref_harness.cpp"), os.path.join(HERE, "ref_limiter.cpp")]
    if bound == "SimpleBound":  # its sources keep the single-device signatures
        srcs.append(os.path.join(HERE, "ref_simplebound.cpp"))
    return srcs


def includes(c):
    _, scheme, _, limiter, bound = c
    dirs = [SRC + "/Geometry", SRC + "/Utils", SRC + "/BoundaryCondition", SRC + "/Base/Euler/Advection",
            SRC + "/Base/Common", SRC + "/Base/Euler/Riemann", SRC + "/Base/Slope_limiter/" + limiter,
            SRC + "/Base/TimeAdvance", SRC + "/BoundaryCondition/" + bound,
            SRC + "/Base/Euler/Advection/" + scheme]
    return ["-I" + d for d in dirs]


# sources g++'s OpenACC front end rejects (an `acc loop` directive before a
# declaration instead of a loop, which nvc++ accepts): compiled without
# -fopenacc, where the directives are ignored and the loops run on the host, as
# the host fallback runs them anyway
NO_ACC = {TEST + "/PNGGeo/Global_Boudary.cpp"}


def build(c):
    out = os.path.join(OUT, lib_name(c))
    srcs = sources(c)
    if os.path.exists(out) and os.path.getmtime(out) >= max(os.path.getmtime(s) for s in srcs):
        return out, "up-to-date"
    defs = ["-DREF_NUM_SCH=%d" % NUM_SCH[c[4]], "-DREF_TMP_W=%d" % (16 if c[1] == "Lax" else 28),
            "-DREF_HAS_SCHEME_INDEX=%d" % (0 if c[4] == "Regular" else 1)]
    objs = []
    for src in [s for s in srcs if s in NO_ACC]:
        obj = out + "." + os.path.basename(src) + ".o"
        r = subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-c", "-w", "-o", obj] + defs +
                           includes(c) + [src], capture_output=True, text=True)
        if r.returncode != 0:
            return out, "FAILED\n" + r.stderr[-3000:]
        objs.append(obj)
    cmd = ["g++", "-O2", "-fopenacc", "-ffp-contract=off", "-fPIC", "-shared", "-w", "-o", out + ".tmp"] + \
        defs + includes(c) + [s for s in srcs if s not in NO_ACC] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return out, "FAILED\n" + r.stderr[-3000:]
    os.replace(out + ".tmp", out)
    for obj in objs:
        os.remove(obj)
    return out, "built"


def build_writedata():
    """The reference's VTK writer alone (Src/Utils/WriteData.cpp) with a C entry point."""
    out = os.path.join(OUT, "libref_writedata.so")
    srcs = [SRC + "/Utils/WriteData.cpp", os.path.join(HERE, "ref_writedata.cpp")]
    if os.path.exists(out) and os.path.getmtime(out) >= max(os.path.getmtime(s) for s in srcs):
        return out, "up-to-date"
    inc = ["-I" + d for d in [SRC + "/Utils", SRC + "/Geometry", SRC + "/Base/Euler/Advection/MUSCL"]]
    # without -fopenacc: g++ rejects its `acc loop` directives outside a compute
    # region (nvc++ accepts them); the loops print on the host either way, and
    # acc_set_device_num comes from libgomp
    cmd = ["g++", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-w", "-o", out + ".tmp"] + inc + srcs + \
        ["-lgomp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return out, "FAILED\n" + r.stderr[-3000:]
    os.replace(out + ".tmp", out)
    return out, "built"


# libpng 1.6.37 (the version the reference's README names): the system's
# runtime library, with the headers of the same release that conda ships (the
# system has no development headers). Linked by path, so nothing else under
# /opt/conda (its older libstdc++) enters the link.
PNG_INC = "/opt/conda/include"
PNG_SO = "/usr/lib/x86_64-linux-gnu/libpng16.so.16"


def build_pngmask():
    """Test/PNGGeo/Level_Set.cpp (the mask level set) with the reference's PNG
    reader (Src/Utils/PngProcess.cpp), ParamReader and Psy_coord, and
    oracle/ref_pngmask.cpp's C entry points, linked against libpng.

    Level_Set.cpp is compiled without -fopenacc: under OpenACC's default data
    rules a scalar referenced in a parallel construct is firstprivate, so g++
    keeps the host's edge count (`edg_num++` under `acc atomic update`,
    Level_Set.cpp:53-77) at 0 and the edge list new double[0] is then overrun
    (:79-95; AddressSanitizer finds the write at :90). Read as serial C++ the
    loops count the edges, which is what the code evidently intends. Its
    scratch arrays are zeroed (operator new[] wrapped to calloc, see
    ref_pngmask.cpp), so the unset border gradients read as no edge."""
    out = os.path.join(OUT, "libref_pngmask.so")
    ls_src = TEST + "/PNGGeo/Level_Set.cpp"
    srcs = [SRC + "/Utils/PngProcess.cpp", SRC + "/Utils/ParamReader.cpp", SRC + "/Geometry/Psy_coord.cpp",
            os.path.join(HERE, "ref_pngmask.cpp")]
    if not os.path.exists(os.path.join(PNG_INC, "png.h")) or not os.path.exists(PNG_SO):
        return out, "FAILED\nlibpng 1.6 (headers under %s, %s) not found" % (PNG_INC, PNG_SO)
    if os.path.exists(out) and os.path.getmtime(out) >= max(os.path.getmtime(s) for s in srcs + [ls_src]):
        return out, "up-to-date"
    inc = ["-I" + d for d in [SRC + "/Geometry", SRC + "/Utils", SRC + "/Base/Euler/Advection/MUSCL", PNG_INC]]
    obj = out + ".Level_Set.o"
    r = subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-c", "-w", "-o", obj] + inc + [ls_src],
                       capture_output=True, text=True)
    if r.returncode != 0:
        return out, "FAILED\n" + r.stderr[-3000:]
    cmd = ["g++", "-O2", "-fopenacc", "-ffp-contract=off", "-fPIC", "-shared", "-w", "-o", out + ".tmp"] + inc + \
        srcs + [obj, PNG_SO, "-Wl,--wrap=_Znam", "-Wl,--wrap=_ZdaPv"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    os.remove(obj)
    if r.returncode != 0:
        return out, "FAILED\n" + r.stderr[-3000:]
    os.replace(out + ".tmp", out)
    return out, "built"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--list", action="store_true")
    a = ap.parse_args()
    if a.list:
        for c in combos():
            print(lib_name(c))
        return 0
    if not os.path.isdir(REF):
        print("build_ref: %s absent; nothing built" % REF)
        return 0
    os.makedirs(OUT, exist_ok=True)
    bad = 0
    with cf.ThreadPoolExecutor(a.j) as ex:
        for out, msg in ex.map(build, combos()):
            if msg.startswith("FAILED"):
                bad += 1
                print(os.path.basename(out), msg)
    for fn in (build_writedata, build_pngmask):
        out, msg = fn()
        if msg.startswith("FAILED"):
            bad += 1
            print(os.path.basename(out), msg)
    print("build_ref: %d libraries, %d failed" % (len(combos()) + 2, bad))
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
