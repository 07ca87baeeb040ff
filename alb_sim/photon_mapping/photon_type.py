from pyalbsim_amd.photon_mapping.photon_type import PhotonType  # noqa: F401
