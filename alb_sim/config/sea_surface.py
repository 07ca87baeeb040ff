from pyalbsim_amd.config.sea_surface import *  # noqa: F401,F403
