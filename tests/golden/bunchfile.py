"""Deterministic bunch file of the `type = file` fixture job (test infrastructure)."""
import numpy as np


def write_bunch_file(path, n, seed):
    """Lab-frame bunch for `type = file` jobs: x y z gbx gby gbz per line, no trailing newline
    (the reader loops on ifstream::good(), src/classes.cpp:379)."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, 100.0, n)
    y = rng.normal(0.0, 100.0, n)
    z = rng.uniform(-5.0, 5.0, n)
    gbx = rng.normal(0.0, 1.0e-3, n)
    gby = rng.normal(0.0, 1.0e-3, n)
    gbz = np.sqrt(100.41 ** 2 - 1.0) + rng.normal(0.0, 0.1, n)
    rows = np.stack([x, y, z, gbx, gby, gbz], axis=1)
    text = "\n".join(" ".join("%.17g" % v for v in r) for r in rows)
    with open(path, "w") as f:
        f.write(text)
