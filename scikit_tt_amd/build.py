"""Build libttals.so in-tree for gfx950 (``python -m scikit_tt_amd.build``)."""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, 'csrc', f) for f in ('ttals_capi.hip', 'ttals_hybrid.hip')]
OUT = os.path.join(HERE, 'libttals.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
         # MFMA results in ordinary VGPRs: accumulation registers (AGPRs) would be
         # allocated on top of the 256 VGPRs and halve the workgroups per CU
         '-mllvm', '-amdgpu-mfma-vgpr-form=1',
         # never emit packed FP32 (v_pk_*_f32) instructions
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']


def _resource_usage(text):
    """Per-kernel {field: value} from hipcc's kernel-resource-usage remarks."""
    usage, name = {}, None
    for line in text.splitlines():
        m = re.search(r'remark:\s+Function Name: (\S+)', line)
        if m:
            name = m.group(1)
            usage[name] = {}
            continue
        m = re.search(r'remark:\s+([A-Za-z ]+?)(?: \[[^]]*\])?: (\d+)', line)
        if m and name:
            usage[name][m.group(1)] = int(m.group(2))
    return usage


def build(verbose=False, out=OUT, defines=()):
    """Compile libttals.so; ``defines`` adds -D options (e.g. TTALS_PROFILE=1
    for a build whose results carry per-phase cycle counters)."""
    tmp = out + '.tmp'
    cmd = [HIPCC] + FLAGS + ['-D' + d for d in defines] + ['-Rpass-analysis=kernel-resource-usage', '-o', tmp] + SRCS
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed building libttals.so')
    os.replace(tmp, out)
    # the batched engine is sized for two resident systems per CU
    for name, u in _resource_usage(r.stderr).items():
        if name.startswith('_Z5k_als') and u.get('Occupancy', 2) < 2:
            sys.stderr.write('warning: %s occupancy %d wave/SIMD (VGPRs %s, AGPRs %s): one system per CU\n'
                             % (name, u['Occupancy'], u.get('VGPRs'), u.get('AGPRs')))
    return out


if __name__ == '__main__':
    print(build(verbose=True))
