"""In-tree build of the native parts (no JIT cache, so the .so files travel with the tree).

  libeio_hip.so                 HIP kernels + C ABI (hipcc, gfx950)
  pyElastic_iso_{float,double}_{nl,born}*.so
                                pybind11 operator modules (C++ host layer over the C ABI)
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EIO_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=" + ARCH,
             # packed FP32 instructions are disabled for all device code on this target
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=CSRC, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_hip():
    out = os.path.join(HERE, "libeio_hip.so")
    _run([HIPCC] + HIP_FLAGS + ["-o", out + ".tmp", "eio_capi.hip"])
    os.replace(out + ".tmp", out)
    return out


def _pybind_flags():
    import pybind11
    inc = [pybind11.get_include(), sysconfig.get_paths()["include"], os.path.join(ROOT, "include"),
           os.path.join(CSRC, "host")]
    # OpenMP: the host-side trace interpolation runs per (shot, component) on all host cores
    return ["-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-fopenmp"] + ["-I" + i for i in inc]


def build_pybind():
    """One extension module per reference module name, each linking libeio_hip.so."""
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    built, jobs = [], []
    srcs = []
    for prec, ctype in (("float", "float"), ("double", "double")):
        for kind in ("nl", "born"):
            name = "pyElastic_iso_%s_%s" % (prec, kind)
            out = os.path.join(HERE, name + suffix)
            cmd = (["g++"] + _pybind_flags() +
                   ["-DEIO_REAL=%s" % ctype, "-DEIO_MODULE=%s" % name,
                    "-DEIO_KIND_%s=1" % kind.upper(), "host/pymodule.cpp"] + srcs +
                   (["-DEIO_DEFINE_INTERP=1"] if kind == "nl" else []) +
                   ["-o", out + ".tmp", "-L" + HERE, "-leio_hip", "-Wl,-rpath,$ORIGIN"])
            jobs.append((cmd, out))
    procs = [(subprocess.Popen(cmd, cwd=CSRC, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True), out)
             for cmd, out in jobs]
    for (p, out), (cmd, _) in zip(procs, jobs):
        txt = p.communicate()[0]
        if p.returncode != 0:
            raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), txt))
        os.replace(out + ".tmp", out)
        built.append(out)
    return built


def build_all(pybind=True):
    libs = [build_hip()]
    if pybind and os.path.exists(os.path.join(CSRC, "host", "pymodule.cpp")):
        libs += build_pybind()
    return libs


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)
