from pyalbsim_amd.config.scene import *  # noqa: F401,F403
