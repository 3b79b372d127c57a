"""Build libttals.so in-tree for gfx950 (``python -m scikit_tt_amd.build``)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'csrc', 'ttals_capi.hip')
OUT = os.path.join(HERE, 'libttals.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
         # never emit packed FP32 (v_pk_*_f32) instructions
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']


def build(verbose=False):
    tmp = OUT + '.tmp'
    cmd = [HIPCC] + FLAGS + ['-o', tmp, SRC]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed building libttals.so')
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose=True))
