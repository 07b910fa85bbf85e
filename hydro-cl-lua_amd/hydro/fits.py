"""FITS images of solver buffers, in the layout the reference writes them.

GridSolver:saveBuffer (hydro/solver/gridsolver.lua:1410-1455) downloads a
buffer of numCells * channels reals (the ghosted grid, channel fastest),
makes an image of width * depth by height with one plane per channel
(x fastest, then z, then y: index i + width (k + depth (j + height ch))) and
saves it as `<basefn>.fits`; GridSolver:save (gridsolver.lua:1465-1470)
names each buffer `<prefix>_<buffer name>`. The file is a primary HDU with
NAXIS1 = width * depth, NAXIS2 = height, NAXIS3 = channels and IEEE reals
(BITPIX -64 for double, -32 for float), big-endian, as the FITS standard
(version 4.0) fixes them. The reference writes through its `image` library
over cfitsio, which is not in this image: the header keywords beyond the
mandatory ones are not compared against it.
"""
import numpy as np

BLOCK = 2880
_BITPIX = {np.dtype(np.float64): -64, np.dtype(np.float32): -32}
_DTYPE = {-64: '>f8', -32: '>f4'}


def _card(key, value):
    if isinstance(value, bool):
        v = 'T' if value else 'F'
    else:
        v = str(int(value))
    return f'{key:<8}= {v:>20}'.ljust(80)


def write_image(path, planes):
    """writes `planes` (channels, height, width) as a FITS primary image"""
    planes = np.asarray(planes)
    if planes.ndim != 3:
        raise ValueError('planes must be (channels, height, width)')
    if planes.dtype not in _BITPIX:
        raise ValueError(f'unsupported dtype {planes.dtype}')
    c, h, w = planes.shape
    cards = [_card('SIMPLE', True), _card('BITPIX', _BITPIX[planes.dtype]), _card('NAXIS', 3),
             _card('NAXIS1', w), _card('NAXIS2', h), _card('NAXIS3', c), 'END'.ljust(80)]
    head = ''.join(cards).encode('ascii')
    head += b' ' * (-len(head) % BLOCK)
    data = np.ascontiguousarray(planes, dtype=_DTYPE[_BITPIX[planes.dtype]]).tobytes()
    data += b'\0' * (-len(data) % BLOCK)
    with open(path, 'wb') as f:
        f.write(head)
        f.write(data)


def read_image(path):
    """(planes (NAXIS3, NAXIS2, NAXIS1) in native byte order, header dict)"""
    raw = open(path, 'rb').read()
    hdr, pos = {}, 0
    while True:
        if pos + BLOCK > len(raw):
            raise ValueError('FITS header without END')
        block = raw[pos:pos + BLOCK].decode('ascii')
        pos += BLOCK
        done = False
        for i in range(0, BLOCK, 80):
            card = block[i:i + 80]
            key = card[:8].strip()
            if key == 'END':
                done = True
                break
            if card[8:10] == '= ':
                v = card[10:].split('/')[0].strip()
                hdr[key] = True if v == 'T' else False if v == 'F' else int(v)
        if done:
            break
    if hdr.get('SIMPLE') is not True or hdr.get('BITPIX') not in _DTYPE:
        raise ValueError('not a FITS image of IEEE reals')
    shape = tuple(hdr[f'NAXIS{n}'] for n in range(hdr['NAXIS'], 0, -1))
    n = int(np.prod(shape))
    dt = np.dtype(_DTYPE[hdr['BITPIX']])
    data = np.frombuffer(raw, dtype=dt, count=n, offset=pos).reshape(shape)
    return data.astype(dt.newbyteorder('=')), hdr


def buffer_planes(U):
    """the reference's image of a ghosted buffer U (z, y, x, channels):
    planes[ch, j, k * width + i] = U[k, j, i, ch]"""
    U = np.asarray(U)
    if U.ndim != 4:
        raise ValueError('buffer must be (z, y, x, channels)')
    d, h, w, c = U.shape
    return np.ascontiguousarray(U.transpose(3, 1, 0, 2)).reshape(c, h, d * w)


def save_buffer(U, basefn):
    """GridSolver:saveBuffer: writes `<basefn>.fits` and returns its name"""
    fn = basefn + '.fits'
    write_image(fn, buffer_planes(U))
    return fn
