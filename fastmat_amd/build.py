"""Build recipe for the fastmat_amd HIP library (gfx950).

Compiles every ``csrc/*.hip`` source into ``fastmat_amd/libfastmat_amd.so``
with hipcc.  Packed-FP32 instruction selection is disabled for device code
(``-packed-fp32-ops``); the host pass of hipcc warns that it does not know the
feature, which is harmless.

Usage: ``python -m fastmat_amd.build`` (also called by ``__graft_entry__.build``).
"""
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfastmat_amd.so")
ARCH = os.environ.get("FASTMAT_AMD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return [
        "--offload-arch=" + ARCH,
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-result",
        "-Wno-unused-command-line-argument",
        "-Wno-unused-value",
    ]


def build(verbose=False, jobs=8):
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [cc] + flags() + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = OUT + ".tmp"
    cmd = [cc] + flags() + ["-shared", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    return OUT


def _drain(procs):
    errors = []
    while procs:
        src, p = procs.pop(0)
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # drop the host pass's notice about the device-only target feature
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l
                 and "is not a recognized feature" not in l]
        if p.returncode != 0:
            errors.append("%s failed:\n%s" % (src, "\n".join(lines)))
        elif any(l.strip() for l in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    if errors:
        raise RuntimeError("\n".join(errors))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
