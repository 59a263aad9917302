"""Shared set-up of oracle-speciated cell states for the GPU parity tests."""
from oracle import speciated_state  # noqa: F401


def host_arrays(cells):
    return {k: (None if v is None else v.copy()) for k, v in cells.arrays.items()}
