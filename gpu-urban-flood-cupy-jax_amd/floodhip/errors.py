"""Error types, mirroring hybrid_flood/operational/errors.py:4-25 of the reference."""


class OperationalError(RuntimeError):
    """Base class for user-facing operational failures."""


class ConfigurationError(OperationalError):
    """The requested model configuration is unsafe or unsupported."""


class BackendUnavailableError(OperationalError):
    """A required compute backend is unavailable."""


class SimulationError(OperationalError):
    """The numerical run failed validation."""


class RainfallInputError(OperationalError):
    """Rainfall input is invalid or incomplete."""


class DataFileMissingError(OperationalError):
    """A required immutable model asset is unavailable."""
