"""Inputs for the qdxt tests: the pixel blocks of a texture's levels, laid
out as mipmapped_texture::qdxt_pack_init does (crnlib/crn_mipmapped_texture.cpp:
2410-2469; crunch2_amd.qdxt.level_blocks)."""
from crunch2_amd import texture
from crunch2_amd.qdxt import level_blocks


def texture_blocks(size, seed, kind=texture.RGB, mips=True, height=None):
    img = texture.synth(size, height or size, seed, kind)
    levels = texture.mip_chain(img) if mips else [img]
    return level_blocks(levels)
