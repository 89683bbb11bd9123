"""Build recipe for libbamf_gpu.so (gfx950) — run in-tree, no JIT cache.

    python -m bamf_amd.build          # compiles bamf_amd/libbamf_gpu.so

Every HIP translation unit is compiled with packed-FP32 instruction
selection disabled (-packed-fp32-ops): ordinary float code is issued as
scalar VALU ops.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libbamf_gpu.so"
HIP_SOURCES = ["runtime.cpp", "mle.hip", "findroi.hip", "rjmcmc.hip", "rjmcmc_host.cpp", "net.hip", "gateway.cpp", "findmap.hip", "tiles.hip", "priors.hip", "psf.hip", "synth.hip"]
ARCH = os.environ.get("BAMF_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libbamf_gpu.so")


def compile_cmds(out: Path = LIB) -> list[list[str]]:
    objs = []
    cmds = []
    build_dir = PKG / "_build"
    build_dir.mkdir(exist_ok=True)
    common = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Wno-unused-command-line-argument", f"-I{ROOT / 'include'}"]
    for src in HIP_SOURCES:
        p = CSRC / src
        if not p.exists():
            continue
        obj = build_dir / (p.stem + ".o")
        cmds.append(common + ["-c", str(p), "-o", str(obj)])
        objs.append(str(obj))
    cmds.append([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(out)] + objs)
    return cmds


def build(verbose: bool = False, jobs: int = 4) -> Path:
    from concurrent.futures import ThreadPoolExecutor

    cmds = compile_cmds()

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        return cmd, subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, r in ex.map(run, cmds[:-1]):
            if r.returncode != 0:
                sys.stderr.write(r.stdout.decode(errors="replace"))
                raise RuntimeError(f"compile failed: {' '.join(cmd)}")
    tmp = LIB.with_suffix(".so.tmp")
    link = [str(tmp) if a == str(LIB) else a for a in cmds[-1]]
    _, r = run(link)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    # the standalone C++ bench harness, linked against the library ($ORIGIN rpath)
    exe = PKG / "bamf_bench"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-Wno-unused-command-line-argument",
           f"-I{ROOT / 'include'}", str(CSRC / "bench_main.cpp"), "-o", str(exe) + ".tmp", f"-L{PKG}", "-lbamf_gpu",
           "-Wl,-rpath,$ORIGIN"]
    _, r = run(cmd)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("bench harness build failed")
    os.replace(str(exe) + ".tmp", exe)
    build_mex(run)
    return LIB


MEX_DIR = PKG / "mex"
MEX_WRAPPERS = ["cRJMCMC_gpu", "findROI_gpu", "gaussMLE_gpu", "findPriors_gpu", "findPSF_gpu"]


def build_mex(run) -> None:
    """The mex wrappers as host shared libraries (g++) against the project's mx
    runtime (csrc/mex/mx), linked to libbamf_gpu.so: bamf_amd/mex/<name>.so,
    loaded by bamf_amd/mexsim.py."""
    MEX_DIR.mkdir(exist_ok=True)
    gxx = shutil.which("g++") or "g++"
    for name in MEX_WRAPPERS:
        out = MEX_DIR / f"{name}.so"
        cmd = [gxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{ROOT / 'include'}",
               f"-I{CSRC / 'mex' / 'mx'}", str(CSRC / "mex" / f"{name}.cpp"), str(CSRC / "mex" / "mx" / "mx.cpp"),
               f"-L{PKG}", "-lbamf_gpu", "-Wl,-rpath,$ORIGIN/..", "-o", str(out) + ".tmp"]
        _, r = run(cmd)
        if r.returncode != 0:
            sys.stderr.write(r.stdout.decode(errors="replace"))
            raise RuntimeError(f"mex wrapper build failed: {name}")
        os.replace(str(out) + ".tmp", out)


if __name__ == "__main__":
    print(build(verbose=True))
