"""Drop-in package: the reference's ``alb_sim`` module paths, backed by pyalbsim_amd."""
