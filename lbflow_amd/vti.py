"""Reader of the drop-in's VTI output (appended raw, UInt64 block headers).

Follows the file layout of the reference writer (solver/src/file/filevti.cpp:39-172)
and the checks of the reference's reader (scripts/vti.py:5-38): every block must
have exactly its piece's length and the pieces must cover the grid once.
"""
import struct
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np


def read_vti(path, pieces=None):
    """Parse an appended-raw VTI with one or more pieces (scripts/vti.py:5-38):
    every DataArray offset must point at a block of exactly its piece's
    length and the pieces must cover the grid exactly once.  Returns
    ((nx, ny, nz), {name: array [nz, ny, nx, n]}); the piece extents
    (x0 x1 y0 y1 z0 z1 each) are appended to the list `pieces` if given."""
    data = Path(path).read_bytes()
    pos = data.index(b'<AppendedData')
    doc = ET.fromstring(data[:pos] + b'</VTKFile>')
    begin = data.index(b'>_', pos) + 2
    im = doc.find('ImageData')
    extent = list(map(int, im.attrib['WholeExtent'].split()))
    nx, ny, nz = [extent[i + 1] - extent[i] + 1 for i in (0, 2, 4)]
    fields = {}
    coverage = np.zeros((nz, ny, nx), np.int64)
    for piece in im.findall('Piece'):
        e = list(map(int, piece.attrib['Extent'].split()))
        if pieces is not None:
            pieces.append(e)
        box = (slice(e[4], e[5] + 1), slice(e[2], e[3] + 1), slice(e[0], e[1] + 1))
        shape = (e[5] - e[4] + 1, e[3] - e[2] + 1, e[1] - e[0] + 1)
        coverage[box] += 1
        for a in piece.find('PointData').findall('DataArray'):
            name = a.attrib['Name']
            n = int(a.attrib.get('NumberOfComponents', '1'))
            off = begin + int(a.attrib['offset'])
            length = struct.unpack_from('<Q', data, off)[0]
            expected = shape[0] * shape[1] * shape[2] * n * 8
            if length != expected:
                raise ValueError(f'{name}: block {length} != {expected}')
            values = np.frombuffer(data, dtype='<f8', count=length // 8, offset=off + 8)
            field = fields.setdefault(name, np.full((nz, ny, nx, n), np.nan))
            field[box] = values.reshape(shape + (n,))
    if (coverage != 1).any():
        raise ValueError('Pieces do not cover whole grid exactly once')
    if any(np.isnan(f).any() for f in fields.values()):
        raise ValueError('Missing field values')
    return (nx, ny, nz), fields
