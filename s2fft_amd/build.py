"""Build the gfx950 shared library ``libs2fft_amd.so`` in-tree with hipcc.

Usage: ``python -m s2fft_amd.build`` (or ``__graft_entry__.build()``).
"""

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libs2fft_amd.so"
ARCH = os.environ.get("S2FFT_AMD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; install ROCm or set HIPCC")


def sources():
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


# Per-source flags.  The network's inference kernel reads every matrix-core
# result with vector instructions (GELU): results in VGPRs instead of AGPRs
# spare one move per element.
EXTRA_FLAGS = {
    "s2h_mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def build(verbose=False):
    """Compile every HIP/C++ source into one shared library (always rebuilds)."""
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    flags = [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        # packed FP32 instructions have produced silent wrong results on this
        # hardware pool: keep the compiler from emitting them
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
        f"-I{HERE.parent / 'include'}",
    ]
    objs = []
    procs = []
    for src in sources():
        obj = objdir / (src.name + ".o")
        cmd = [hipcc(), *flags, *EXTRA_FLAGS.get(src.name, []), "-c", str(src), "-o", str(obj)]
        if src.suffix == ".cpp":
            cmd.insert(1, "-x")
            cmd.insert(2, "hip")
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = False
    for src, proc in procs:
        out, _ = proc.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature: harmless
        text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
        if proc.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {src.name} failed\n{text}\n")
        elif verbose and text.strip():
            sys.stderr.write(f"--- {src.name}\n{text}\n")
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB.with_suffix(".so.tmp")
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
