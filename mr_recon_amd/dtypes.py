"""Default dtypes of the hot path (reference: dtypes.py:5-6): float32 / complex64."""

import numpy as np
import torch

real_dtype = torch.float32
complex_dtype = torch.complex64
np_real_dtype = np.float32
np_complex_dtype = np.complex64
