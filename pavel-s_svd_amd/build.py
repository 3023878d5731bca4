"""Build the gfx950 kernel library and its CPython binding in-tree.

* ``libpavelsvd.so`` -- every ``csrc/*.hip`` file, compiled by hipcc for gfx950 only; exports the
  C ABI declared in ``include/pavelsvd.h``.
* ``_hip<EXT_SUFFIX>`` -- ``csrc/capi.cpp`` compiled by g++, linked against ``libpavelsvd.so``
  with an ``$ORIGIN`` rpath, so the package directory is self-contained.

Both land in this package directory.  Run ``python pavel-s_svd_amd/build.py`` (or call
``build()``); objects go to ``_build/`` beside this file.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / 'csrc'
INCLUDE = ROOT / 'include'
OBJ = PKG / '_build'
ROCM = Path(os.environ.get('ROCM_PATH', '/opt/rocm'))
HIPCC = str(ROCM / 'bin' / 'hipcc')
LIB = PKG / 'libpavelsvd.so'
EXT = PKG / ('_hip' + sysconfig.get_config_var('EXT_SUFFIX'))

HIP_FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-Wall', '-Wno-unused-result',
             # packed FP32 instructions are kept out of every kernel
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-I', str(INCLUDE), '-I', str(CSRC)]


def _run(cmd, verbose):
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'build step failed ({res.returncode}):\n{" ".join(cmd)}\n{res.stdout}')
    return res.stdout


def _compile_hip(src, verbose):
    obj = OBJ / (src.stem + '.o')
    _run([HIPCC, *HIP_FLAGS, '-c', str(src), '-o', str(obj)], verbose)
    return obj


def build(verbose=False, jobs=None):
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob('*.hip'))
    jobs = jobs or min(8, os.cpu_count() or 1, len(sources))
    with ThreadPoolExecutor(jobs) as pool:
        objs = list(pool.map(lambda s: _compile_hip(s, verbose), sources))
    tmp = LIB.with_suffix('.so.tmp')
    _run([HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', *map(str, objs), '-o', str(tmp)], verbose)
    os.replace(tmp, LIB)
    py_inc = sysconfig.get_paths()['include']
    tmp = Path(str(EXT) + '.tmp')
    _run(['g++', '-O2', '-std=c++17', '-fPIC', '-shared', '-I', py_inc, '-I', str(INCLUDE),
          str(CSRC / 'capi.cpp'), '-L', str(PKG), '-lpavelsvd', "-Wl,-rpath,$ORIGIN", '-o', str(tmp)],
         verbose)
    os.replace(tmp, EXT)
    return LIB, EXT


if __name__ == '__main__':
    for path in build(verbose='-v' in sys.argv):
        print(path)
