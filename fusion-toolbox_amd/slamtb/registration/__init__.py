"""Registration algorithms (slamtb/registration/__init__.py)."""
from .multiscale import MultiscaleRegistration
from .result import RegistrationResult, RegistrationVerifier
from .icp import ICPOdometry
from .correspondence_map import CorrespondenceMap
