from nozomi_amd.substrate import createBoundaryZ, create_pad  # noqa: F401
