"""MI355X-native GPU-driven frame of vk_gltf_viewer: meshlet cull + compaction,
Hi-Z pyramid, visibility-buffer rasterization and material resolve, as HIP
kernels for gfx950 behind the C ABI of include/vkgv.h."""
