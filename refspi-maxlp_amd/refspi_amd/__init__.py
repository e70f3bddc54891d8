"""MI355X-native (gfx950) reconstruction hot path of RefSPI-MaxLP.

Modules: ``emc`` (EMC E-step facade), ``maxlp`` (MaxLP object update), ``dset``
(sparse-photon dataset), ``synth`` (seeded synthetic frames), ``h5io`` (photons/output
HDF5 files), ``_lib`` (ctypes binding of the C-ABI), ``build`` (hipcc build).
"""
__all__ = ['emc', 'maxlp', 'dset', 'synth', 'h5io', 'build']
