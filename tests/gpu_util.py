"""Shared helpers of the GPU parity tests."""
import numpy as np


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def flush(msg):
    print(msg, flush=True)
