"""Synthetic physical inputs shared by the tests (Haar unitaries and
threshold-detector O matrices, reference sampling/gbs.py:241-275); the
generators live in gbskernels_amd.workloads."""
from gbskernels_amd.workloads import haar_unitary, threshold_O  # noqa: F401
