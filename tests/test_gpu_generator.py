"""The generator's device stages (ncsm_inputs_build_gpu): index lists,
instructions and shapes identical to the host generator's
(ncsm_inputs_build, itself pinned by brute force and by the compiled
reference accepting its files), matrix elements equal to the last units in
the place (the device's log and cos against glibc's)."""
import time

import numpy as np
import pytest

from jupiterncsm_amd.inputs import SYSTEMS, SystemConfig, build

pytestmark = pytest.mark.gpu

CASES = {
    "he4_nmax4": SYSTEMS["he4_nmax4"],
    "he4_nmax6": SYSTEMS["he4_nmax6"],
    "odd_3p2n_nmax3_3b": SystemConfig(3, 2, 3, twice_m=1, max_body=3, seed=4, sigma3=0.2),
    "li6_nmax2_3b": SystemConfig(3, 3, 2, max_body=3, seed=11, sigma3=0.1),
    "no_one_body_4p4n_nmax2": SystemConfig(4, 4, 2, max_body=3, include_one_body=False, seed=2),
    "li6_nmax6_3b": SYSTEMS["li6_nmax6_3b"],
}


def generated(cfg, device):
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    inp = build(cfg, device=device)
    return inp, time.perf_counter() - t


@pytest.mark.parametrize("name", list(CASES))
def test_device_generator_matches_host(name):
    print(f"case generator {name}", flush=True)
    cfg = CASES[name]
    host, t_host = generated(cfg, False)
    dev, t_dev = generated(cfg, True)
    print(f"{name}: D = {host.dimension}, {host.list_offsets[-1]} list entries, "
          f"{host.matrix_offsets[-1]} elements; host {t_host:.2f} s, device {t_dev:.2f} s",
          flush=True)
    assert dev.dimension == host.dimension
    for field in ("blocks", "list_offsets", "list_entries", "list_settings", "matrix_offsets",
                  "matrix_shapes", "orbitals"):
        a, b = getattr(dev, field), getattr(host, field)
        assert a.shape == b.shape and np.array_equal(a, b), field
    assert np.array_equal(dev.instructions, host.instructions)
    a, b = dev.matrix_elements, host.matrix_elements
    assert a.shape == b.shape
    scale = np.abs(b).max()
    diff = np.abs(a - b)
    print(f"elements: {np.mean(a == b):.4f} bit-identical, max |diff| {diff.max():.2e} "
          f"(max |v| {scale:.2e})", flush=True)
    assert diff.max() <= 1e-14 * scale
    assert np.all(diff <= 8 * np.spacing(np.abs(b)) + 1e-300)
