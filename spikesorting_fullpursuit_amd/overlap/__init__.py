from .binary_pursuit_parallel import binary_pursuit  # noqa: F401
