"""Build the HIP library in-tree: stride_amd/_lib/libstride_amd.so (gfx950).

Run ``python -m stride_amd.build``. Compiles each .hip file under csrc/ with
hipcc for gfx950 only and links one shared library exposing the C ABI declared
in include/stride_amd.h.
"""

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT_DIR = os.path.join(HERE, '_lib')
LIB = os.path.join(OUT_DIR, 'libstride_amd.so')
ARCH = 'gfx950'

# packed FP32 instructions are disabled for all device code
FLAGS = ['-O3', '--offload-arch=%s' % ARCH, '-fPIC', '-std=c++17', '-Wall', '-Wno-unused-result',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I', os.path.join(HERE, '..', 'include')]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def _headers():
    inc = os.path.join(HERE, '..', 'include')
    return [os.path.join(d, f) for d in (CSRC, inc) for f in os.listdir(d) if f.endswith('.h')]


def build(verbose=False, objdir=None):
    """Compile every source and link the library. With ``objdir``, objects are kept there and
    a source is recompiled only when it or a header is newer than its object."""
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    os.makedirs(OUT_DIR, exist_ok=True)
    keep = objdir is not None
    objdir = objdir or OUT_DIR
    os.makedirs(objdir, exist_ok=True)
    newest_header = max(os.path.getmtime(h) for h in _headers())
    objs = []
    jobs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src)[:-4] + '.o')
        fresh = keep and os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(src), newest_header)
        if not fresh:
            cmd = [hipcc, '-c', src, '-o', obj] + FLAGS
            if verbose:
                print(' '.join(cmd), flush=True)
            jobs.append(cmd)
        objs.append(obj)
    # one compiler process per source, at most one per host core
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 16))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for done in pool.map(lambda c: subprocess.run(c), jobs):
            if done.returncode != 0:
                raise subprocess.CalledProcessError(done.returncode, done.args)
    tmp = LIB + '.tmp'
    cmd = [hipcc, '-shared', '-o', tmp] + objs + ['--offload-arch=%s' % ARCH, '-fPIC']
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if not keep:
        for obj in objs:
            os.remove(obj)
    return LIB


if __name__ == '__main__':
    # --objdir DIR: keep objects in DIR and rebuild only what changed
    objdir = sys.argv[sys.argv.index('--objdir') + 1] if '--objdir' in sys.argv else None
    print(build(verbose='-v' in sys.argv, objdir=objdir))
