"""Build libkcore.so (HIP, gfx950) and the drop-in executables in-tree.

    python -m pico_amd.build          # library + executables
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIBDIR, "libkcore.so")

ARCH = os.environ.get("KC_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 instructions stay off in every HIP translation unit.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-fvisibility=hidden", "-Wno-unused-command-line-argument", "-Wno-unused-value",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIB_SOURCES = [
    os.path.join(CSRC, "kc_peel.hip"),
    os.path.join(CSRC, "kc_compact.hip"),
    os.path.join(CSRC, "kc_hindex.hip"),
    os.path.join(CSRC, "kc_prep.hip"),
    os.path.join(CSRC, "kc_batch.hip"),
    os.path.join(CSRC, "kc_gen.hip"),
    os.path.join(CSRC, "kc_dyn.hip"),
    os.path.join(CSRC, "kc_capi.cpp"),
    os.path.join(CSRC, "kc_batch_capi.cpp"),
    os.path.join(CSRC, "kc_dyn_capi.cpp"),
    os.path.join(CSRC, "kc_server.cpp"),
    os.path.join(CSRC, "kc_noalloc.cpp"),
    os.path.join(HOST, "kc_host.cpp"),
    os.path.join(HOST, "kc_cache.cpp"),
]

CLI_TOOLS = ["gpp", "peelone", "peelone_dyn", "HistoCore", "CntCore", "NbrCore"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r.stdout


def build_lib(jobs=8):
    """Compile each translation unit to an object in parallel, then link."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(ROOT, "_build", "obj")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in LIB_SOURCES:
        if not os.path.exists(src):
            continue
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith((".hip", ".cpp")) else []
        cmd = [HIPCC] + HIP_FLAGS + lang + ["-c", src, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_cli():
    """Host C++ drop-ins for the reference executables, linked to libkcore."""
    src = os.path.join(HOST, "kc_cli.cpp")
    if not os.path.exists(src):
        return []
    os.makedirs(BINDIR, exist_ok=True)
    exe = os.path.join(BINDIR, "kcore_cli")
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src,
          "-o", exe, "-L" + LIBDIR, "-lkcore", "-Wl,-rpath,$ORIGIN/../lib"])
    out = []
    for name in CLI_TOOLS:
        link = os.path.join(BINDIR, name)
        if os.path.lexists(link):
            os.remove(link)
        os.symlink("kcore_cli", link)
        out.append(link)
    return out


def build_all():
    build_lib()
    build_cli()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
