"""MI355X-native six-face GAMERA-Cube step: host-side package."""
from . import config  # noqa: F401
from . import problem  # noqa: F401
