"""Build the HIP engine library (kmos_amd/libkmcgpu.so) for gfx950."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "kmcg_engine.hip")
DEPS = [SRC, os.path.join(HERE, "csrc", "kmcg_device.h"), os.path.join(HERE, "csrc", "kmcg_otf.hip"),
        os.path.join(REPO, "include", "kmcg.h"), os.path.join(REPO, "include", "kmcg_otf.h")]
LIB = os.path.join(HERE, "libkmcgpu.so")
# the same source with the step loop's phase counters compiled in
PHASES_LIB = os.path.join(HERE, "libkmcgpu_phases.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    # the reference's x86-64 build has no FMA: keep every FP64 op separate
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"),
]


def build(force=True, verbose=False):
    """Compile the engine and its phase-profiling variant (in parallel);
    always recompiles unless force=False and both libraries are newer than
    their sources.  Returns the engine's path."""
    jobs = []
    for lib, extra in ((LIB, []), (PHASES_LIB, ["-DKMCG_PHASES=1"])):
        if not force and os.path.exists(lib) and \
                all(os.path.getmtime(lib) >= os.path.getmtime(d) for d in DEPS):
            continue
        tmp = lib + ".tmp%d" % os.getpid()
        cmd = [HIPCC] + FLAGS + extra + [SRC, "-o", tmp]
        jobs.append((lib, tmp, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                                text=True)))
    failed = []
    for lib, tmp, proc in jobs:
        out = proc.communicate()[0]
        if proc.returncode != 0:
            failed.append(out)
            continue
        if verbose:
            print(out)
        os.replace(tmp, lib)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(failed))
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
