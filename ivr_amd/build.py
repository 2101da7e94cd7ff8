"""Build recipe of libivr.so (hipcc for gfx950, gcc for the host C).

Objects go to ivr_amd/_build/, the library to ivr_amd/libivr.so.  Every HIP
source is compiled with contraction off (fused multiply-adds are explicit
fmaf calls, matching the CPU restatement bit for bit) and with packed fp32
instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libivr.so")
COMPAT_LIB = os.path.join(HERE, "libivr_compat.so")
SERVER = os.path.join(HERE, "ivrserver")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("IVR_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include"),
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-Wall",
           "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def sources():
    """(hip, c) sources of libivr.so; compat* files form libivr_compat.so"""
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    c = sorted(f for f in os.listdir(CSRC) if f.endswith(".c"))
    return hip, c


def _split(files):
    return ([f for f in files if not f.startswith(("compat", "ivrserver"))],
            [f for f in files if f.startswith("compat")])


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    hip, c = sources()
    jobs = []
    for f in hip:
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, f), "-o",
                                           os.path.join(OUT, f + ".o")])
    for f in c:
        if f.startswith("ivrserver"):
            continue
        jobs.append(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, f), "-o",
                                         os.path.join(OUT, f + ".o")])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    core_h, compat_h = _split(hip)
    core_c, compat_c = _split(c)
    objs = [os.path.join(OUT, f + ".o") for f in core_h + core_c]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, LIB)
    # drop-in library: reference names, host globals resolved at load time
    objs = [os.path.join(OUT, f + ".o") for f in compat_h + compat_c]
    tmp = COMPAT_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L" + HERE, "-livr", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(tmp, COMPAT_LIB)
    # steering server executable over the C-ABI
    tmp = SERVER + ".tmp"
    _run(["gcc"] + C_FLAGS + [os.path.join(CSRC, "ivrserver.c"), "-o", tmp, "-L" + HERE, "-livr",
                              "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(tmp, SERVER)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
