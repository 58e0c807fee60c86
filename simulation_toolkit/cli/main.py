from nozomi_amd.cli.main import main, run_simulation  # noqa: F401

if __name__ == "__main__":
    import sys
    sys.exit(main())
