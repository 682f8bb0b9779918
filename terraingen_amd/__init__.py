# SPDX-License-Identifier: BSD-3-Clause
"""TerrainGen's noise + crater shaping hot path on MI355X (gfx950): the hip backend's host side.

``native`` binds the C ABI of ``libtg_hip.so`` (kernels) and ``libterraingen_amd.so`` (placement,
preparation, binning, the backend registry and the tile server); ``pipeline`` composes terrains and
tile batches from them; ``serving`` drives the tile server with a camera's request stream.
"""
