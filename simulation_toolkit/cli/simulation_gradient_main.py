from nozomi_amd.cli.drivers import simulation_gradient_main  # noqa: F401
