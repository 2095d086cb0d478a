"""Builds libsil_hip.so (the HIP C-ABI library) in-tree for gfx950.

    python -m similie_amd.build            # compile
    python -m similie_amd.build --check    # compile only if sources are newer
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libsil_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SIL_OFFLOAD_ARCH", "gfx950")

SOURCES = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    # packed-FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-lambda-capture",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def needs_build():
    if not LIB.exists():
        return True
    t = LIB.stat().st_mtime
    deps = SOURCES + list(CSRC.glob("*.hpp")) + list((ROOT / "include").glob("*.h"))
    return any(p.stat().st_mtime > t for p in deps)


CPP_TESTS = ROOT / "tests" / "cpp"


def build_cpp_tests(verbose=False):
    """C++ parity tests of include/sil/exterior.hpp, linked against the library."""
    src = CPP_TESTS / "exterior_tests.cpp"
    out = CPP_TESTS / "exterior_tests"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++20", f"-I{ROOT / 'include'}", str(src),
           f"-L{PKG}", "-lsil_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../similie_amd",
           "-o", str(out)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=str(ROOT))
    probe = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++20", "-fPIC", "-shared", f"-I{CSRC}",
             str(CPP_TESTS / "mfma_probe.hip"), "-o", str(CPP_TESTS / "libmfma_probe.so")]
    subprocess.run(probe, check=True, cwd=str(ROOT))
    return out


def build(force=True, verbose=False):
    """Compiles every source to an object in parallel (in a temporary
    directory), then links libsil_hip.so."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    if not force and not needs_build():
        return LIB
    compile_flags = [f for f in FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="sil_build_") as tmpdir:
        objs = [Path(tmpdir) / (src.name + ".o") for src in SOURCES]

        def compile_one(pair):
            src, obj = pair
            cmd = [HIPCC, *compile_flags, "-c", str(src), "-o", str(obj)]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True, cwd=str(ROOT))

        jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            list(pool.map(compile_one, zip(SOURCES, objs)))
        tmp = LIB.with_suffix(".so.tmp")
        link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
        if verbose:
            print(" ".join(link), flush=True)
        subprocess.run(link, check=True, cwd=str(ROOT))
    os.replace(tmp, LIB)
    build_cpp_tests(verbose)
    return LIB


if __name__ == "__main__":
    build(force="--check" not in sys.argv, verbose=True)
    print(f"built {LIB}")
