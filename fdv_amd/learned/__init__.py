"""The learned motion proposer: its training (train.py) and the rate-distortion
evaluation against the exhaustive search (bdrate.py).  The trained weights ship
as proposer.f32 (fp32, the layout of fdv_amd/csrc/fdvgpu_nn.hip)."""
