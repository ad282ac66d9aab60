"""MI355X-native MPIDForce (multipoles + induced dipoles, PME) evaluation."""
from .mpidforce import MPIDForce, MPIDForceError  # noqa: F401
from .serialization import XmlSerializer  # noqa: F401
from .forcefield import ForceField, PDBFile  # noqa: F401
