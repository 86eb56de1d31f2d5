"""In-tree build of the native solver library for gfx950 (hipcc) and of the Node-API addon.

`build()` always recompiles: librnozzle.so (HIP kernels + C ABI + host setup) into
build_out/, and, when node headers are present, build_out/nozzle_napi.node.
"""
import os
import shutil
import subprocess

from ._lib import LIB_PATH, ROOT

REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["kernels.hip", "instance.hip", "runtime.hip", "batch.hip", "cutcell.hip"]
# per-file flags: the cut-cell geometry must round exactly as the reference's JavaScript
# (FP64 with float32 stores), so no multiply-add contraction there
FILE_FLAGS = {"cutcell.hip": ["-ffp-contract=off"]}
CPP_SOURCES = ["setup.cpp", "errors.cpp"]

# device code: no packed-FP32 instructions (they are avoided on this platform); small
# per-thread arrays stay in private memory rather than being moved into LDS, whose size
# sets how many stage workgroups share a CU
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-disable-promote-alloca-to-lds"]
# the stage kernels' arguments (instance table, instance count, XCD chunk, tile table) preloaded into
# SGPRs at wave launch, so the first table read does not wait on a kernel-argument load
STAGE_FLAGS = ["-mllvm", "-amdgpu-kernarg-preload-count=4"]


def stage_shapes():
    """(name, type, TX, TY, W) of every entry of csrc/stage_shapes.h, in order."""
    import re
    text = open(os.path.join(CSRC, "stage_shapes.h")).read()
    return [(m.group(1), m.group(2), int(m.group(3)), int(m.group(4)), int(m.group(5)))
            for m in re.finditer(r"X\((\w+),\s*(float|double),\s*(\d+),\s*(\d+),\s*(\d+)\)", text)]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout[-6000:])
    return r.stdout


def build_lib(jobs=8):
    out_dir = os.path.dirname(LIB_PATH)
    obj_dir = os.path.join(out_dir, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", "-I", INCLUDE, "-I", CSRC]
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-c", *common, *DEVICE_FLAGS, "-munsafe-fp-atomics",
               *FILE_FLAGS.get(src, []), os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    # one object per tile shape of the stage kernels
    for name, typ, tx, ty, w in stage_shapes():
        obj = os.path.join(obj_dir, f"stage_{name}.o")
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-c", *common, *DEVICE_FLAGS, *STAGE_FLAGS, "-munsafe-fp-atomics",
               f"-DNZ_STAGE_NAME={name}", f"-DNZ_STAGE_R={typ}", f"-DNZ_STAGE_TX={tx}", f"-DNZ_STAGE_TY={ty}",
               f"-DNZ_STAGE_W={w}",
               os.path.join(CSRC, "stage_inst.hip"), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        cmd = ["g++", "-c", *common, "-D__HIP_PLATFORM_AMD__", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out[-6000:])
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_napi():
    """Node-API addon binding the C ABI for the JavaScript drop-in (js/gpu_solver.js)."""
    node_inc = "/usr/include/node"
    if not os.path.exists(os.path.join(node_inc, "node_api.h")):
        return None
    out = os.path.join(os.path.dirname(LIB_PATH), "nozzle_napi.node")
    src = os.path.join(ROOT, "js", "nozzle_napi.cpp")
    if not os.path.exists(src):
        return None
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", node_inc, "-I", INCLUDE, src,
          "-L", os.path.dirname(LIB_PATH), "-lrnozzle", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build():
    lib = build_lib()
    napi = build_napi()
    return lib, napi


def clean():
    shutil.rmtree(os.path.dirname(LIB_PATH), ignore_errors=True)
