"""Test-side alias of the synthetic model builders."""
from seismic_amd.synthetic import (ricker, layered_anomaly, smooth_random, faulted_layers,  # noqa: F401
                                   gaussian_smooth, gradient_anomalies)
