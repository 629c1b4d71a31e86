"""Raw file exchange with the HamiltonFastMarching (HFM) executables
(agd/Eikonal/FileIO.py:14-64), so that inputs and outputs of this solver can
be compared with a CPU HFM run where one is available.

A dictionary is stored as a pair of files:
  * ``<prefix>_Format.txt``: for every entry, its key on one line, then
    ``0`` for a number, ``-1`` followed by the (unicode-escaped) string for a
    string, or the number of dimensions followed by one line per dimension
    for an array; each entry ends with an empty line;
  * ``<prefix>_Data.dat``: the numbers and the arrays (row-major), in entry
    order, as raw float64.
This is the layout the HFM executables read and write.
"""
import numbers
import os
import platform
import subprocess

import numpy as np

from ._device import host_array


def RawToFiles(params, prefix='input'):
    """Writes the dictionary `params` (numbers, strings, arrays) to
    ``prefix_Format.txt`` / ``prefix_Data.dat``."""
    if not isinstance(prefix, str):
        raise TypeError("prefix must be a string")
    lines, chunks = [], []
    for key, val in params.items():
        if isinstance(val, str):
            lines += [key, '-1', val.encode('unicode_escape').decode('utf-8'), '']
            continue
        if isinstance(val, numbers.Number):
            lines += [key, '0', '']
            chunks.append(np.array([val], dtype=np.float64))
            continue
        if isinstance(val, np.ndarray) or hasattr(val, '__array__') or hasattr(val, 'get'):
            arr = np.asarray(host_array(val))
            lines += [key, str(arr.ndim)] + [str(n) for n in arr.shape] + ['']
            chunks.append(arr.astype(np.float64).ravel())
            continue
        raise ValueError(f"Invalid type for key {key} : {type(val)}")
    with open(prefix + '_Format.txt', 'w') as f:
        f.write(''.join(line + '\n' for line in lines))
    data = np.concatenate(chunks) if chunks else np.zeros(0, np.float64)
    data.astype('d').tofile(prefix + '_Data.dat')


def FilesToRaw(prefix='output'):
    """Reads ``prefix_Format.txt`` / ``prefix_Data.dat`` back into a
    dictionary (numbers as float64 scalars, arrays as float64 arrays)."""
    data = np.fromfile(prefix + '_Data.dat', dtype='d')
    out = {}
    pos = 0
    with open(prefix + '_Format.txt') as f:
        while True:
            key = f.readline().strip()
            if not key:
                break
            kind = int(f.readline())
            if kind == -1:
                out[key] = f.readline().strip().encode('utf-8').decode('unicode_escape')
            elif kind == 0:
                out[key] = data[pos]
                pos += 1
            else:
                dims = [int(f.readline()) for _ in range(kind)]
                size = int(np.prod(dims))
                out[key] = data[pos:pos + size].reshape(dims)
                pos += size
            f.readline()  # the empty line ending the entry
    return out


def WriteCallRead(inputData, executable, binary_dir='', working_dir=None,
                  inputPrefix="input", outputPrefix="output"):
    """Writes inputData, runs an HFM executable on it and reads its output
    (agd/Eikonal/FileIO.py:66-86)."""
    if working_dir is None:
        working_dir = binary_dir
    if platform.system() == 'Linux' and binary_dir == '.':
        binary_dir = ''
    RawToFiles(inputData, os.path.join(working_dir, inputPrefix))
    proc = subprocess.Popen([os.path.join(binary_dir, executable), inputPrefix, outputPrefix],
                            stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                            universal_newlines=True, cwd=working_dir or None)
    for line in iter(proc.stdout.readline, ""):
        print(line, end='')
    ret = proc.wait()
    if ret != 0:
        print('Returned with exit code ', ret)
    outputData = FilesToRaw(os.path.join(working_dir, outputPrefix))
    outputData['retcode'] = ret
    return outputData
