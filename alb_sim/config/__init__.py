from pyalbsim_amd.config import *  # noqa: F401,F403
