"""Build the gfx950 shared library (libgpu_dsm.so) and the gpu_DSM executable in-tree.

hipcc flags that matter for results:
  -ffp-contract=off   no implicit FMA contraction (the FP32 operation order is part of
                      the specification the CPU restatement reproduces);
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division / square root;
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 VALU ops.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libgpu_dsm.so")
EXE = os.path.join(OUT, "gpu_DSM")

ARCH = os.environ.get("DSM_OFFLOAD_ARCH", "gfx950")
COMMON = [
    "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["dsm_kernels.hip", "dsm_engine.cpp", "dsm_driver.cpp", "dsm_serve.cpp", "dsm_noalloc.hip"]
GEOMS = [0, 1, 2, 3, 4, 5, 6, 7]  # KMC chain geometries (one translation unit each; 6 = bins 0 and 1 merged:
                                  # the row form, then the row form with two register groups or the
                                  # pair form; 7 = pair form with two register groups)


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; ROCm is required to build gpu_dsm_amd")
    return exe


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False, jobs=None):
    os.makedirs(OUT, exist_ok=True)
    units = [(os.path.join(CSRC, s), os.path.join(OUT, os.path.splitext(s)[0] + ".o"), []) for s in SOURCES]
    units += [(os.path.join(CSRC, "dsm_kmc_inst.hip"), os.path.join(OUT, "dsm_kmc_g%d.o" % g), ["-DDSM_GEOM=%d" % g])
              for g in GEOMS]
    # largest geometries first: they dominate the compile time
    units.sort(key=lambda u: -int(u[2][0].split("=")[1]) if u[2] else 0)
    jobs = jobs or max(1, min(len(units), os.cpu_count() or 1, 16))
    pending, running, objs = list(units), [], []
    while pending or running:
        while pending and len(running) < jobs:
            src, obj, extra = pending.pop(0)
            cmd = [hipcc(), "-x", "hip", *COMMON, *DEVICE, *extra, "-c", src, "-o", obj]
            p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            running.append((p, cmd, obj))
        p, cmd, obj = running.pop(0)
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            for q, _, _ in running:
                q.kill()
            raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc(), *DEVICE, "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    tmpx = EXE + ".tmp"
    _run([hipcc(), "-x", "hip", *COMMON, *DEVICE, os.path.join(CSRC, "gpu_DSM.cpp"), "-o", tmpx,
          "-L", OUT, "-lgpu_dsm", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmpx, EXE)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
