"""Builds blml_amd/libblml_hip.so for gfx950 with hipcc (in-tree, so it travels with the source tree).
    python band-limited-ml_amd/build.py"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, 'csrc', 'engine.hip')
OUT = os.path.join(HERE, 'blml_amd', 'libblml_hip.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I', os.path.join(ROOT, 'include'), '-I', os.path.join(HERE, 'csrc')]


def build(verbose=True):
    tmp = OUT + '.tmp'
    cmd = [HIPCC] + FLAGS + [SRC, '-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    build()
    sys.exit(0)
