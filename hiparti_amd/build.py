"""Build libHiParTI.so in-tree (hiparti_amd/lib/libHiParTI.so).

Host sources (csrc/host/*.cpp) are compiled with g++, HIP sources
(csrc/device/*.hip) with hipcc for gfx950 only.  The MTTKRP kernel
instantiations (csrc/device/mttkrp_inst.hip) are compiled once per tensor
order 2..8 so they build in parallel.  Objects go to hiparti_amd/_build/.

Usage:  python -m hiparti_amd.build      (or build() from Python)
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libHiParTI.so")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("PTI_OFFLOAD_ARCH", "gfx950")
# Packed-FP32 instructions stay disabled for all device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-pthread", "-Wall", "-Wno-unused-function"]
ORDERS = range(2, 9)


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in p.stdout.splitlines()
                    if "packed-fp32-ops" not in line and line.strip())
    if p.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (p.returncode, " ".join(cmd), out))
    return out


def _jobs():
    host = []
    for f in sorted(os.listdir(os.path.join(CSRC, "host"))):
        if f.endswith(".cpp"):
            src = os.path.join(CSRC, "host", f)
            obj = os.path.join(BUILD, "host_" + f[:-4] + ".o")
            host.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
    dev = []
    for f in sorted(os.listdir(os.path.join(CSRC, "device"))):
        if not f.endswith(".hip"):
            continue
        src = os.path.join(CSRC, "device", f)
        if f == "mttkrp_inst.hip":
            for n in ORDERS:
                obj = os.path.join(BUILD, "dev_mttkrp_inst_n%d.o" % n)
                dev.append([HIPCC] + HIP_FLAGS + ["-DPTI_NMODES=%d" % n, "-c", src, "-o", obj])
        else:
            obj = os.path.join(BUILD, "dev_" + f[:-4] + ".o")
            dev.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    return host + dev


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    cmds = _jobs()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out:
                print(out)
    objs = [c[c.index("-o") + 1] for c in cmds]
    tmp = LIB + ".tmp"
    # Only the C ABI is exported (exports.map); the C++ internals stay local.
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-pthread", "-Wl,--version-script=" + os.path.join(CSRC, "exports.map")])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
