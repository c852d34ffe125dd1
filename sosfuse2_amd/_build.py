"""Build libsosfuse2_hip.so in-tree with hipcc for gfx950 (the .so ships beside the sources)."""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / 'csrc'
LIB = HERE / 'libsosfuse2_hip.so'
SOURCES = ('transport.hip', 'sh.hip', 'ground.hip', 'project.hip', 'iterate.hip', 'geometry.hip', 'capi.cpp', 'mlp.hip')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = [
    '--offload-arch=gfx950', '-O3', '-fPIC', '-std=c++17', '-fgpu-flush-denormals-to-zero',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wno-unused-result',
]


def _compile(src, obj):
    lang = ['-x', 'hip'] if src.suffix == '.cpp' else []
    cmd = [HIPCC, *FLAGS, *lang, '-c', str(src), '-o', str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'E_BUILD: {" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    return obj


def build(verbose=True):
    """Compile every HIP source for gfx950 and link the shared library. Always recompiles."""
    subprocess.run([sys.executable, str(CSRC / 'gen_constants.py'), str(CSRC / 'sh_constants.h')], check=True)
    out = CSRC / '_obj'
    out.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    with ThreadPoolExecutor(min(8, len(srcs))) as pool:
        objs = list(pool.map(lambda s: _compile(s, out / (s.stem + '.o')), srcs))
    tmp = LIB.with_suffix('.so.tmp')
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', *map(str, objs), '-o', str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'E_BUILD: link failed\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print(f'built {LIB} ({LIB.stat().st_size // 1024} KiB)')
    return LIB


if __name__ == '__main__':
    build()
