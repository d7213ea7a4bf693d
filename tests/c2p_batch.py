from asterx_amd.zones import make_batch  # noqa: F401
