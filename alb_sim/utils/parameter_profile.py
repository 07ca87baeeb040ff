from pyalbsim_amd.utils.parameter_profile import *  # noqa: F401,F403
