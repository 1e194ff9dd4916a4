"""Parameter vectors of a generated scene for the differentiable entry points."""
import numpy as np


def scene_params(scene):
    """albedo (M,3), emission (M,3), intensity (L,), texels (level 0 of every
    texture, RGB / 255) of a SheenScene, as float32 arrays."""
    s = scene.contents
    M, L = s.num_materials, s.num_lights
    alb = np.array([[s.materials[i].diffuse_color.x, s.materials[i].diffuse_color.y, s.materials[i].diffuse_color.z]
                    for i in range(M)], np.float32)
    em = np.array([[s.materials[i].emission_color.x * s.materials[i].emission,
                    s.materials[i].emission_color.y * s.materials[i].emission,
                    s.materials[i].emission_color.z * s.materials[i].emission] for i in range(M)], np.float32)
    inten = np.array([s.lights[i].intensity for i in range(L)], np.float32)
    tex = []
    for t in range(s.num_textures):
        w, h = s.texture_widths[t], s.texture_heights[t]
        if not s.texture_pixels[t] or w <= 0 or h <= 0:
            continue
        px = np.ctypeslib.as_array(s.texture_pixels[t], shape=(w * h * 3,))
        tex.append(px.astype(np.float32) / np.float32(255.0))
    texels = np.concatenate(tex).astype(np.float32) if tex else np.zeros(0, np.float32)
    return {"albedo": alb, "emission": em, "intensity": inten, "texels": texels}


def loss_delta(a_plus, a_minus, target):
    """sum((a+ - t)^2) - sum((a- - t)^2) over RGB, in float64 without the
    cancellation of two large sums."""
    ap = a_plus[..., :3].astype(np.float64)
    am = a_minus[..., :3].astype(np.float64)
    t = target[..., :3].astype(np.float64)
    return float(np.sum((ap - am) * (ap + am - 2.0 * t)))
