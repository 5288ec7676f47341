"""Host-side checks of the CNN batched workload (no GPU)."""
import numpy as np


def test_layer_table_and_sparsity(srt):
    layers = srt.cnn.cnn_layers(seed=0)
    assert len(layers) == 16
    sizes = {64, 128, 256, 512, 1024, 2048}
    for l in layers:
        assert l.M in sizes and l.K in sizes and l.hw in (3136, 784, 196, 49)
        assert 0.80 <= l.sparsity <= 0.98
        A = l.A.to_dense().numpy()
        assert A.shape == (l.M, l.K)
        density = np.count_nonzero(A) / A.size
        # Bernoulli(1 - s) per entry: within 6 standard deviations
        sd = np.sqrt((1 - l.sparsity) * l.sparsity / A.size)
        assert abs(density - (1 - l.sparsity)) <= 6 * sd + 1.0 / A.size
    # bottleneck: every reduce layer (M < K) is followed in its group by an expand (M > K)
    for g in range(4):
        grp = layers[4 * g:4 * g + 4]
        assert len({l.hw for l in grp}) == 1
        assert any(l.M > l.K for l in grp) and any(l.M < l.K for l in grp)
    again = srt.cnn.cnn_layers(seed=0)
    for a, b in zip(layers, again):
        assert np.array_equal(a.A.to_dense().numpy(), b.A.to_dense().numpy())
    other = srt.cnn.cnn_layers(seed=1)
    assert not np.array_equal(layers[0].A.to_dense().numpy(), other[0].A.to_dense().numpy())


def test_traffic_and_flops(srt):
    layers = srt.cnn.cnn_layers(seed=0)[:3]
    batch = 2
    want_b = want_f = 0
    for l in layers:
        A = l.A.to_dense().numpy()
        n = l.hw * batch
        used = int(np.count_nonzero(np.any(A != 0, axis=0)))
        nnz = int(np.count_nonzero(A))
        want_b += used * n * 4 + l.M * n * 4 + nnz * 8
        want_f += 2 * nnz * n
    assert srt.cnn.algorithmic_bytes(layers, batch) == want_b
    assert srt.cnn.flops(layers, batch) == want_f
