#!/usr/bin/env python3
"""Build libyell_amd.so (HIP for gfx950) and the oracle libraries.

    python3 yell_amd/build.py            # full rebuild
    python3 yell_amd/build.py --changed  # rebuild objects older than their inputs

Every translation unit is compiled by hipcc in parallel into yell_amd/_build/,
then linked into yell_amd/lib/libyell_amd.so.  The oracle (test
infrastructure) is built by `make -C oracle`.
"""
import argparse
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libyell_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("YA_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    # keep hipcc from forming packed-FP32 instructions (there is no float code
    # on this path, the flag is a guard)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def headers():
    return glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(CSRC, "*.inc")) + \
        glob.glob(os.path.join(ROOT, "include", "*.h"))


def obj_for(src):
    return os.path.join(OBJ, os.path.basename(src) + ".o")


def stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def compile_one(src):
    cmd = [HIPCC] + CFLAGS + ["-c", "-x", "hip", src, "-o", obj_for(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    return src, r.returncode, msg


def build_lib(changed_only=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    hdrs = headers()
    todo = [s for s in sources() if not changed_only or stale(obj_for(s), [s] + hdrs)]
    jobs = jobs or min(16, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, msg in ex.map(compile_one, todo):
            if msg.strip():
                print(msg, file=sys.stderr)
            if rc != 0:
                raise RuntimeError("hipcc failed on %s" % src)
    objs = [obj_for(s) for s in sources()]
    if todo or stale(LIB, objs):
        cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB] + objs + ["-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    return LIB


GMP_INC = "/opt/conda/include"
SYS_GMP = "/usr/lib/x86_64-linux-gnu/libgmp.so.10"


def build_cpp_tests():
    """tests/cpp/*.cpp -> tests/cpp/build/<name>, linked against libyell_amd.so;
    test_shim.cpp is also built with -DYELL_USE_32BITS_MODULI (test_shim32)."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "build")
    os.makedirs(out_dir, exist_ok=True)
    have_gmp = os.path.exists(os.path.join(GMP_INC, "gmp.h")) and os.path.exists(SYS_GMP)
    for src in sorted(glob.glob(os.path.join(src_dir, "*.cpp"))):
        name = os.path.splitext(os.path.basename(src))[0]
        variants = [(name, [])]
        if name == "test_shim":
            variants.append((name + "32", ["-DYELL_USE_32BITS_MODULI"]))
        for exe_name, defs in variants:
            exe = os.path.join(out_dir, exe_name)
            cmd = ["g++", "-std=c++17", "-O2", "-Wall"] + defs + ["-I", os.path.join(ROOT, "include"), src,
                                                                "-o", exe, "-L", os.path.dirname(LIB), "-lyell_amd",
                                                                "-Wl,-rpath,$ORIGIN/../../../yell_amd/lib"]
            if have_gmp and not defs:
                # poly2mpz of the C++ drop-in: GMP headers here, the system library by path
                i = cmd.index(src)
                cmd[i:i] = ["-I", GMP_INC]
                cmd.append(SYS_GMP)
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("C++ test build failed:\n" + r.stdout + r.stderr)


def build_oracle():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--changed", action="store_true", help="only rebuild stale objects")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    build_lib(a.changed, a.j)
    build_cpp_tests()
    if not a.no_oracle:
        build_oracle()
    print(LIB)


if __name__ == "__main__":
    main()
