"""Exception types with the reference's names and bases
(crowddynamics/exceptions.py:10-30)."""


class CrowdDynamicsException(Exception):
    """Base exception."""


class InvalidType(CrowdDynamicsException, TypeError):
    """Arguments had an invalid type (e.g. an unknown agent record type)."""


class InvalidValue(CrowdDynamicsException, ValueError):
    """Arguments had an invalid value."""


class DeviceError(CrowdDynamicsException, RuntimeError):
    """The HIP library reported an error or is unavailable."""
