"""The synthetic workload generators (mcgpu.configs): the torch form of the
splitmix64 uniform generator equals the numpy form bit for bit (on the CPU
here; tests/test_parity_gpu.py repeats it on the device), and the seed chain
equals the oracle's serial rnd_seed chain."""
import numpy as np
import pytest


@pytest.mark.parametrize("seed,offset", [(2003, 0), (4004, 123_456_789), (5004, 7)])
def test_uniform_torch_equals_numpy_on_cpu(seed, offset):
    from mcgpu import configs
    a = configs.uniform(seed, 200_000, offset)
    b = configs.uniform_torch(seed, 200_000, offset, device="cpu").numpy()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_seed_chain_equals_oracle(oracle):
    from mcgpu import configs
    assert np.array_equal(configs.seed_chain(50_000), oracle.rnd_seed_chain(configs.ROOT_SEED, 50_000))
    assert np.array_equal(configs.seed_chain(1000, start=4000),
                          oracle.rnd_seed_chain(configs.ROOT_SEED, 5000)[4000:])


def test_adjoint1000_batch_layout():
    from mcgpu import configs
    slabs, sa, ss, g, total = configs.adjoint1000_batch(4, 1000)
    w = configs.adjoint1000(10)
    assert total == 4000 and len(sa) == 4000 and g.sum() == 400
    assert np.array_equal(sa[:1000], w.sigma_a) and np.array_equal(ss[:1000], w.sigma_s)
    assert [s[4] for s in slabs] == [0, 1000, 2000, 3000]
    assert [s[5] for s in slabs] == [0, 1000, 2000, 3000]
