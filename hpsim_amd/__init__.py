"""MI355X-native HPSim tracking engine.

``hpsim_amd.HPSim`` is the drop-in for the reference extension module
``HPSim`` (Beam, BeamLine, SpaceCharge, Simulator); ``hpsim_amd.hpsim`` is the
user-unit facade of the reference pylib/hpsim.py; ``hpsim_amd.configs`` builds
the synthetic benchmark beams and lattices.
"""
__all__ = ["HPSim", "hpsim", "configs"]
