"""FITS dumps of solver buffers (gridsolver.lua:1410-1470) on the CPU: the
reference's plane layout (x fastest, then z, then y, one plane per channel),
the FITS structure the standard fixes (80-character cards in 2880-byte
blocks, big-endian IEEE reals, data padded to a block), and a write/read
round trip. The reference writes through cfitsio, which is absent here: its
optional header keywords are not compared (parity unpinned beyond the
mandatory cards and the data)."""
import numpy as np
import pytest

from hydro import fits


def test_plane_layout_matches_reference_index():
    rng = np.random.default_rng(0)
    d, h, w, c = 3, 5, 4, 6
    U = rng.standard_normal((d, h, w, c))
    planes = fits.buffer_planes(U)
    flat = planes.ravel()
    for ch in range(c):
        for k in range(d):
            for j in range(h):
                for i in range(w):
                    # getIndex(ch, i, j, k) = i + width (k + depth (j + height ch))
                    assert flat[i + w * (k + d * (j + h * ch))] == U[k, j, i, ch]


@pytest.mark.parametrize('dtype,bitpix', [(np.float64, -64), (np.float32, -32)])
def test_file_structure_and_round_trip(tmp_path, dtype, bitpix):
    rng = np.random.default_rng(1)
    U = rng.standard_normal((1, 7, 9, 6)).astype(dtype)
    fn = fits.save_buffer(U, str(tmp_path / 'results_UBuf'))
    assert fn.endswith('results_UBuf.fits')
    raw = open(fn, 'rb').read()
    assert len(raw) % 2880 == 0
    cards = [raw[i:i + 80].decode('ascii') for i in range(0, 2880, 80)]
    assert cards[0] == 'SIMPLE  = ' + 'T'.rjust(20) + ' ' * 50
    assert cards[1].startswith('BITPIX  = ') and int(cards[1][10:30]) == bitpix
    assert [int(cards[k][10:30]) for k in range(2, 6)] == [3, 9, 7, 6]   # NAXIS, NAXIS1 = w * d, NAXIS2 = h, NAXIS3
    assert cards[6].rstrip() == 'END'
    n = U.size * np.dtype(dtype).itemsize
    assert raw[2880:2880 + n] == fits.buffer_planes(U).astype(np.dtype(dtype).newbyteorder('>')).tobytes()
    assert set(raw[2880 + n:]) <= {0}
    planes, hdr = fits.read_image(fn)
    assert hdr['NAXIS'] == 3 and planes.dtype == dtype
    assert np.array_equal(planes, fits.buffer_planes(U))
    # the reference's reader sees channel ch of cell (x, y) at planes[ch, y, x]
    assert planes[0, 3, 2] == U[0, 3, 2, 0] and planes[5, 6, 8] == U[0, 6, 8, 5]


def test_rejects_bad_shapes(tmp_path):
    with pytest.raises(ValueError):
        fits.buffer_planes(np.zeros((3, 4, 5)))
    with pytest.raises(ValueError):
        fits.write_image(str(tmp_path / 'x.fits'), np.zeros((2, 3, 4), dtype=np.int32))
