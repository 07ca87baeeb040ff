from pyalbsim_amd.photon_mapping.photon_map_data import PhotonMapData, load_photon_maps, save_photon_maps  # noqa: F401
