"""Builds the Fortran drivers (placeholder until the drivers land)."""


def build(verbose: bool = False) -> None:
    return None
