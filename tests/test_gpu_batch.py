"""Batches of scenes in one wavefront pipeline (sheen_batch_*): every scene
of a batch renders as it does alone, for any chunking, with or without the
captured hipGraph."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# uneven sizes (not multiples of the 8x8 tile), spp 1..8, 1 or 2 lights
SIZES = [(64, 64, 1), (120, 72, 3), (40, 88, 8), (96, 96, 2), (33, 17, 5), (128, 64, 4), (72, 72, 6), (16, 16, 1)]


@pytest.fixture(scope="module")
def batch():
    from sheen_amd import scenes
    from sheen_amd.batch import Batch
    sc = [scenes.generate(4, seed=scenes.BATCH_SEED + k, detail=(1500 if k % 2 else 12000)) for k in range(6)]
    sc.append(scenes.generate(1))    # 64 spheres + plane, one light
    sc.append(scenes.generate(101))  # meshes, cubes, glass
    b = Batch(sc, SIZES)
    yield {"b": b, "scenes": sc}
    b.close()


def test_batched_images_equal_serial_renders(batch):
    b = batch["b"]
    s_serial = b.render_serial()
    s_batch = b.render()
    for k in range(b.n):
        assert np.array_equal(b.image(k), b.image(k, serial=True)), k
    for key in ("samples", "primary_rays", "shadow_rays", "reflection_rays", "refraction_rays", "total_rays"):
        assert s_batch[key] == s_serial[key], key


def test_graph_replays_match_direct_launches(batch):
    b = batch["b"]
    b.set_param("graph", 0)
    b.render()
    ref = [b.image(k) for k in range(b.n)]
    b.set_param("graph", 1)
    for _ in range(3):
        b.render_serial()  # other work on the device between replays
        b.render()
        for k in range(b.n):
            assert np.array_equal(b.image(k), ref[k]), k


def test_chunk_beyond_the_slot_index_range_is_refused(batch):
    """A chunk-local slot index travels in 28 bits of a secondary-ray entry:
    larger chunks are refused before any buffer is sized for them."""
    b = batch["b"]
    chunks = b.chunks
    with pytest.raises(RuntimeError, match="chunk above 2\\^28 - 1 slots"):
        b.set_param("chunk", 1 << 28)
    with pytest.raises(RuntimeError, match="chunk above"):
        b.set_param("chunk", 0x7FFFFFFF)
    assert b.chunks == chunks
    b.render()


def test_chunks_across_scene_boundaries(batch):
    b = batch["b"]
    b.set_param("chunk", 0)
    b.render()
    ref = [b.image(k) for k in range(b.n)]
    b.set_param("chunk", 64 * 8 * 7)  # 7 tiles of 8 spp: chunks straddle scenes
    assert b.chunks > b.n
    b.render()
    for k in range(b.n):
        assert np.array_equal(b.image(k), ref[k]), k
    b.set_param("chunk", 0)


def test_sorted_secondary_rays_in_a_batch(batch):
    """Sorting each round's secondary rays (their scenes mixed in one queue)
    changes no batched image, with chunks inside and across scenes."""
    b = batch["b"]
    b.set_param("sec_sort", 0)
    b.render()
    ref = [b.image(k) for k in range(b.n)]
    for mode, chunk in ((1, 0), (2, 0), (2, 64 * 8 * 7)):
        b.set_param("sec_sort", mode)
        b.set_param("chunk", chunk)
        b.render()
        for k in range(b.n):
            assert np.array_equal(b.image(k), ref[k]), (mode, chunk, k)
    b.set_param("sec_sort", 0)
    b.set_param("chunk", 0)


def test_scene_table_sizes(batch):
    """The slot -> scene search gives the same images whether the scenes'
    first slots all sit in the LDS table or only every step-th of them."""
    b = batch["b"]
    b.render()
    ref = [b.image(k) for k in range(b.n)]
    for cap in (1, 3, 7):  # steps 8, 3, 2 over the 8 scenes
        b.set_param("scene_table", cap)
        b.render()
        for k in range(b.n):
            assert np.array_equal(b.image(k), ref[k]), (cap, k)
    b.set_param("scene_table", 512)
    with pytest.raises(RuntimeError):
        b.set_param("scene_table", 513)


def test_batched_scene_matches_oracle(batch):
    from oracle import Oracle
    b = batch["b"]
    b.render()
    k = 6  # the 64-sphere scene
    w, h, a = SIZES[k]
    ref = Oracle(batch["scenes"][k].scene, threads=min(16, os.cpu_count() or 1)).render(w, h, a)["image"]
    img = b.image(k)
    rel = np.abs(img[..., :3] - ref[..., :3]) / np.maximum(np.abs(ref[..., :3]), 1e-3)
    assert rel.max() <= 1e-4


def test_bad_render_params_are_refused():
    from sheen_amd import scenes
    from sheen_amd.batch import Batch
    sc = scenes.generate(1)
    with pytest.raises(RuntimeError, match="bad render params for batch scene 1"):
        Batch([sc, sc], [(32, 32, 1), (32, 32, 65)])


def test_batch_of_every_primitive_kind_media_beams_and_caustics():
    """Scenes with analytic primitives, CSG, hair, media, spot beams and their
    own photon maps batch with plain ones: every batched image equals the
    scene's own render bit for bit, under the pipeline built for every kind,
    with and without the captured graph, and one of them equals the oracle."""
    from oracle import Oracle
    from sheen_amd import scenes
    from sheen_amd.batch import Batch
    sc = [scenes.generate(104), scenes.generate(105), scenes.generate(103),
          scenes.load_sheen(os.path.join(os.path.dirname(__file__), "scenes", "media.sheen")),
          scenes.generate(4, seed=scenes.BATCH_SEED + 3, detail=3000), scenes.generate(1)]
    sizes = [(96, 54, 2), (80, 60, 3), (64, 48, 2), (72, 56, 2), (64, 64, 4), (48, 40, 1)]
    with Batch(sc, sizes) as b:
        sb = b.render()
        ss = b.render_serial()
        for key in ("samples", "primary_rays", "shadow_rays", "reflection_rays", "refraction_rays", "total_rays"):
            assert sb[key] == ss[key], key
        ref = [b.image(k, serial=True) for k in range(b.n)]
        for k in range(b.n):
            assert np.array_equal(b.image(k), ref[k]), k
        b.set_param("graph", 0)
        b.render()
        for k in range(b.n):
            assert np.array_equal(b.image(k), ref[k]), k
        W, H, A = sizes[1]
        o = Oracle(sc[1].scene, threads=min(16, os.cpu_count() or 1)).render(W, H, A)["image"]
        rel = np.abs(ref[1][..., :3] - o[..., :3]) / np.maximum(np.abs(o[..., :3]), 1e-3)
        assert rel.max() <= 1e-4, rel.max()
