from pyalbsim_amd.photon_mapping.photon_map_data import PhotonMapData  # noqa: F401
