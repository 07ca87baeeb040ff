from pyalbsim_amd.config.sensor import *  # noqa: F401,F403
