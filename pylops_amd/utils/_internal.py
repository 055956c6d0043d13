"""Small argument helpers (pylops/utils/_internal.py:9-40)."""
from collections.abc import Sized

import numpy as np


def _value_or_sized_to_array(value_or_sized, repeat=1):
    return (np.asarray(value_or_sized) if isinstance(value_or_sized, Sized)
            else np.array([value_or_sized] * repeat))


def _value_or_sized_to_tuple(value_or_sized, repeat=1):
    return (tuple(value_or_sized) if isinstance(value_or_sized, Sized)
            else tuple([value_or_sized] * repeat))
