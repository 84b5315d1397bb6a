"""Single-cell operators and kets (reference constants.py:3-17)."""
import numpy as np

KET_0 = np.array([1.0, 0.0])
KET_1 = np.array([0.0, 1.0])
PROJECTION_KET_0 = np.array([[1.0, 0.0], [0.0, 0.0]])
PROJECTION_KET_1 = np.array([[0.0, 0.0], [0.0, 1.0]])
S_OPERATOR = np.array([[0.0, 1.0], [1.0, 0.0]])  # lowering + raising = sigma^x
IDENTITY = np.eye(2)
