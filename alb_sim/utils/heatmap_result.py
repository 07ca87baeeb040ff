from pyalbsim_amd.utils.heatmap_result import *  # noqa: F401,F403
