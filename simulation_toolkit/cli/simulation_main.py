from nozomi_amd.cli.drivers import simulation_main  # noqa: F401
