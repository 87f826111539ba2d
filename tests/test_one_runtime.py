"""Loading the library before torch must not put a second HIP runtime in
the process: after native.lib() and `import torch`, exactly one
libamdhip64 is mapped."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
from rover_amd import native
native.lib()
import torch
maps = set()
for line in open('/proc/self/maps'):
    path = line.split()[-1]
    if 'libamdhip64' in path:
        maps.add(path)
print(len(maps))
"""


def test_library_first_maps_one_hip_runtime():
    out = subprocess.run([sys.executable, "-c", PROBE], capture_output=True, text=True,
                         timeout=300, check=True, cwd=ROOT)
    assert out.stdout.strip().splitlines()[-1] == "1", out.stdout
