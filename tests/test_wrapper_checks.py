"""CPU checks of the Python wrapper's argument validation: tensors of the
wrong shape or too-small buffers are refused on the host, before any
kernel could read or write past them."""
import pytest

from radeonrays_sdk_amd import api


def test_ray_tensors_must_be_r_by_8_float32():
    import torch

    assert api._check_rays(torch.zeros((5, 8), dtype=torch.float32), "t") == 5
    assert api._check_rays(torch.zeros((0, 8), dtype=torch.float32), "t") == 0
    for bad in (torch.zeros(40, dtype=torch.float32),            # flattened records
                torch.zeros((5, 4), dtype=torch.float32),         # half records
                torch.zeros((5, 8), dtype=torch.float64),         # wrong element size
                torch.zeros((8, 5), dtype=torch.float32).t()):    # not contiguous
        with pytest.raises(ValueError):
            api._check_rays(bad, "t")


def test_buffers_smaller_than_needed_are_refused():
    import torch

    api._check_size(torch.zeros(64, dtype=torch.uint8), 64, "hits", "t")
    with pytest.raises(ValueError):
        api._check_size(torch.zeros(63, dtype=torch.uint8), 64, "hits", "t")
