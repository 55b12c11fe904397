"""The reference-format benchmark (benchmark/benchmark.h:55-97) run on the GPU
under test: the harness's size list, its normalised rate
(transforms/s * max(1, N ln N) * 1e-6, benchmark.h:80-81) and its CSV/JS
files, for the per-call and the batched path."""
import math

import pytest

from fft_amd import benchmark_csv as bc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,mode", [("f64", "single"), ("f32", "batched")])
def test_measure_and_write(tmp_path, precision, mode):
    rows = bc.measure(precision, mode, max_size=4096, min_ms=2.0)
    assert [r[0] for r in rows] == bc.reference_sizes(4096)
    for n, rate, scaled in rows:
        assert rate > 0 and math.isfinite(rate), (n, rate)
        assert scaled == pytest.approx(rate * max(1.0, n * math.log(n)) * 1e-6, rel=1e-12)
    prefix = str(tmp_path / f"sfft-{precision}-{mode}")
    bc.write(rows, "GPU", prefix)
    lines = open(prefix + ".csv").read().splitlines()
    assert lines[0] == "size,ops/sec,GPU"
    assert len(lines) == len(rows) + 1
    for line, (n, rate, scaled) in zip(lines[1:], rows):
        f = line.split(",")
        assert int(f[0]) == n
        assert float(f[1]) == pytest.approx(rate, rel=1e-12)
        assert float(f[2]) == pytest.approx(scaled, rel=1e-12)
    js = open(prefix + ".js").read()
    assert js.startswith('addResults("GPU", [') and js.endswith("]);")
    assert js.count("{size: ") == len(rows)
