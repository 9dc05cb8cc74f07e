"""Build the HIP library in-tree: exahype_amd/libexahype_amd.so (gfx950)."""
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libexahype_amd.so")
SOURCES = ["csrc/aderdg_kernels.hip", "csrc/capi.hip", "csrc/limiter_kernels.hip", "csrc/limiter_capi.hip",
           "csrc/patchset_capi.hip", "csrc/fv_kernels.hip", "csrc/fv_capi.hip", "csrc/linear_kernels.hip", "csrc/amr_kernels.hip",
           "csrc/amr_capi.hip", "csrc/output_kernels.hip", "csrc/predictor_mfma.hip",
           "csrc/corrector_kernels.hip"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-shared",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + os.path.join(ROOT, "include", "exahype_amd", "device"),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


RESOURCES = os.path.join(HERE, "kernel_resources.txt")


def parse_resource_remarks(text):
    """Per-kernel resource usage from hipcc's -Rpass-analysis=kernel-resource-usage remarks."""
    kernels = {}
    name = None
    for line in text.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            kernels[name] = {}
            continue
        m = re.search(r"remark:\s+([A-Za-z ]+?)(?: \[[^\]]*\])?: (\d+)", line)
        if m and name:
            kernels[name][m.group(1).strip()] = int(m.group(2))
    return kernels


# Per-source flags.  The matrix-core predictor keeps the MFMA accumulators in
# ordinary VGPRs (the AGPR form copies them back and forth and needs more
# registers in total).
EXTRA_FLAGS = {"predictor_mfma.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src, obj):
    cmd = [hipcc()] + [f for f in HIPCC_FLAGS if f != "-shared"] + EXTRA_FLAGS.get(os.path.basename(src), []) + [
        "-c", "-Rpass-analysis=kernel-resource-usage", "-o", obj, src]
    res = subprocess.run(cmd, capture_output=True, text=True)
    return src, cmd, res


def build(verbose=False):
    """Compile for gfx950, one object per source in parallel (every device
    symbol is local to its source file), then link.  Fails if any kernel
    spills to scratch memory: a spill stored inside a divergent region saves
    only the active lanes, so kernels must stay within their register budget."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    objs = [os.path.join(objdir, os.path.basename(s)[:-4] + ".o") for s in SOURCES]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    remarks = []
    with ThreadPoolExecutor(jobs) as ex:
        for src, cmd, res in ex.map(lambda so: _compile(*so), zip(srcs, objs)):
            if verbose:
                print(" ".join(cmd), flush=True)
            if res.returncode != 0:
                sys.stderr.write(res.stdout + res.stderr)
                raise RuntimeError(f"hipcc failed on {os.path.basename(src)}")
            remarks.append(res.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("linking libexahype_amd.so failed")
    kernels = parse_resource_remarks("\n".join(remarks))
    with open(RESOURCES, "w") as f:
        for k in sorted(kernels):
            r = kernels[k]
            f.write(f"{k} vgpr={r.get('VGPRs')} agpr={r.get('AGPRs')} scratch={r.get('ScratchSize')} "
                    f"occupancy={r.get('Occupancy')} lds={r.get('LDS Size')}\n")
    spilling = sorted(k for k, r in kernels.items() if r.get("ScratchSize", 0) > 0)
    if not kernels:
        raise RuntimeError("no kernel resource remarks parsed from hipcc output")
    if spilling:
        raise RuntimeError("kernels spill to scratch memory:\n  " + "\n  ".join(spilling))
    os.replace(tmp, LIB)
    build_user_example(verbose)
    return LIB


USER_EXAMPLE = os.path.join(HERE, "libexa_user_example.so")


def build_user_example(verbose=False):
    """A user's PDE compiled into the device kernels through
    include/exahype_amd/user_solver.hpp, as a user would build it: its own
    shared object, independent of the library's."""
    src = os.path.join(HERE, "examples", "user_euler.hip")
    cmd = [hipcc()] + HIPCC_FLAGS + ["-Rpass-analysis=kernel-resource-usage", "-o", USER_EXAMPLE + ".tmp", src]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("building the user-solver example failed")
    kernels = parse_resource_remarks(res.stderr)
    spilling = sorted(k for k, r in kernels.items() if r.get("ScratchSize", 0) > 0)
    if spilling:
        raise RuntimeError("user-solver kernels spill to scratch memory:\n  " + "\n  ".join(spilling))
    os.replace(USER_EXAMPLE + ".tmp", USER_EXAMPLE)
    return USER_EXAMPLE


if __name__ == "__main__":
    print(build(verbose=True))
