from pyalbsim_amd.config.heatmap import *  # noqa: F401,F403
