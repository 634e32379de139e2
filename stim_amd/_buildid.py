"""Source hash of the native library: SHA-256 over the HIP/C++ sources and the C-ABI header.

build.sh embeds it in libstim_amd.so (mfs_build_id), and the package compares it with the tree it is
imported from, so a run can show that the library it loaded was built from these sources.
Runs without third-party imports (build.sh calls it with plain python3).
"""

import hashlib
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
_SUFFIXES = ('.hip', '.cc', '.h', '.sh')


def source_files():
    csrc = os.path.join(_HERE, 'csrc')
    files = [os.path.join('stim_amd', 'csrc', f) for f in sorted(os.listdir(csrc)) if f.endswith(_SUFFIXES)]
    files.append(os.path.join('include', 'stim_amd.h'))
    return files


def source_hash():
    h = hashlib.sha256()
    for rel in source_files():
        with open(os.path.join(_ROOT, rel), 'rb') as f:
            data = f.read()
        h.update(rel.replace(os.sep, '/').encode() + b'\0')
        h.update(len(data).to_bytes(8, 'little'))
        h.update(data)
    return h.hexdigest()[:16]


if __name__ == '__main__':
    print(source_hash())
