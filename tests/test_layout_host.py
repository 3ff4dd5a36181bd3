"""The padded array layout on the host (no kernel runs): empty_field's views index like the dense
arrays, put row k = 1 of every column on a 128-byte line, leave a full pitch of storage past every
column, and field_pitch / require_fields tell the layouts apart."""
import pytest
import torch

from torchannel_amd import _lib as L


@pytest.mark.parametrize("dtype,nz,pitch", [(torch.float64, 128, 144), (torch.float64, 130, 144),
                                            (torch.float32, 128, 160), (torch.float32, 256, 288),
                                            (torch.float64, 14, 16), (torch.float32, 30, 32)])
def test_padded_pitch(dtype, nz, pitch):
    assert L.padded_pitch(nz, dtype) == pitch


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape", [(17, 14, 18), (18, 13, 18), (18, 14, 17), (5, 6, 130)])
def test_empty_field_layout(dtype, shape):
    P = L.padded_pitch(shape[2], dtype)
    f = L.empty_field(shape, P, dtype, "cpu")
    esz = f.element_size()
    assert tuple(f.shape) == shape and f.stride() == (shape[1] * P, P, 1)
    assert (f.data_ptr() + esz) % 128 == 0 and (P * esz) % 128 == 0
    # storage covers planes * columns * pitch elements from the first entry (flat passes)
    assert f.untyped_storage().nbytes() >= f.data_ptr() - f.untyped_storage().data_ptr() + \
        shape[0] * shape[1] * P * esz
    assert bool((f == 0).all())
    assert L.field_pitch(f) == P
    src = torch.arange(shape[0] * shape[1] * shape[2], dtype=dtype).reshape(shape)
    f.copy_(src)
    assert torch.equal(f, src) and f[3 % shape[0], 1, 2] == src[3 % shape[0], 1, 2]
    d = L.empty_field(shape, 0, dtype, "cpu")
    assert d.is_contiguous() and L.field_pitch(d) == 0
    c = L.clone_field(f)
    assert L.field_pitch(c) == P and torch.equal(c, src)
    assert L.field_like(d).is_contiguous()


def test_field_pitch_rejects_other_layouts():
    x = torch.zeros(6, 7, 8, dtype=torch.float64)
    with pytest.raises(RuntimeError):
        L.field_pitch(x.transpose(0, 1))
    with pytest.raises(RuntimeError):
        L.field_pitch(x[:, :, ::2])


def test_empty_field_rejects_bad_pitch():
    with pytest.raises(ValueError):
        L.empty_field((4, 4, 20), 16, torch.float64, "cpu")      # shorter than the column
    with pytest.raises(ValueError):
        L.empty_field((4, 4, 20), 24, torch.float64, "cpu")      # not whole 128-byte lines
