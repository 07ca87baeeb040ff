from pyalbsim_amd.config.run import *  # noqa: F401,F403
