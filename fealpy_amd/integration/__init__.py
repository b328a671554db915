"""Bindings a fealpy maintainer would add to route fealpy's own calls into
libfealpy_amd.so (see fealpy_backend)."""
