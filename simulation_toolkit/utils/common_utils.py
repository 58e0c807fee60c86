from nozomi_amd.substrate import (  # noqa: F401
    import_array_geometry_full_path, save_data_array_to_pickle, split_matrix_to_list)
