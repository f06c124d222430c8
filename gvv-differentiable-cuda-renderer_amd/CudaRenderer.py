"""Drop-in replacement for the reference's ``CudaRenderer`` module (python/CudaRenderer.py).

Same class, constructor keywords and defaults, getters and output layouts; tensors are
PyTorch-ROCm tensors on the current HIP device instead of TensorFlow tensors, and gradients
flow through ``torch.autograd`` instead of ``tf.GradientTape``.  The ``...TF`` getter names are
kept so caller code ports unchanged; ``...OpenCV`` getters return BGR numpy arrays as the
reference's cv2 conversions did.

    import CudaRenderer
    renderer = CudaRenderer.CudaRendererGpu(faces_attr=..., texCoords_attr=..., ...,
                                            vertexPos_input=vpos, ..., intrinsics_input=K)
    loss = ((renderer.getRenderBufferTF() - target) ** 2).sum()
    loss.backward()
"""
from __future__ import annotations

import numpy as np
import torch

from gvv_amd import cuda_renderer_gpu, cuda_renderer_grad_gpu  # noqa: F401  (op-level API, as customOperators)


class CudaRendererGpu:
    """python/CudaRenderer.py:30-164: the constructor runs the forward op."""

    def __init__(self,
                 faces_attr=[],
                 texCoords_attr=[],
                 numberOfVertices_attr=-1,
                 numberOfCameras_attr=-1,
                 renderResolutionU_attr=-1,
                 renderResolutionV_attr=-1,
                 albedoMode_attr='textured',
                 shadingMode_attr='shaded',
                 image_filter_size_attr=1,
                 texture_filter_size_attr=1,
                 compute_normal_map_attr=False,

                 vertexPos_input=None,
                 vertexColor_input=None,
                 texture_input=None,
                 shCoeff_input=None,
                 targetImage_input=None,
                 extrinsics_input=[],
                 intrinsics_input=[],

                 nodeName='CudaRenderer'):

        self.faces_attr = faces_attr
        self.texCoords_attr = texCoords_attr
        self.numberOfVertices_attr = numberOfVertices_attr
        self.numberOfCameras_attr = numberOfCameras_attr
        self.renderResolutionU_attr = renderResolutionU_attr
        self.renderResolutionV_attr = renderResolutionV_attr
        self.albedoMode_attr = albedoMode_attr
        self.shadingMode_attr = shadingMode_attr
        self.image_filter_size_attr = image_filter_size_attr
        self.texture_filter_size_attr = texture_filter_size_attr
        self.compute_normal_map_attr = compute_normal_map_attr

        self.vertexPos_input = vertexPos_input
        self.vertexColor_input = vertexColor_input
        self.texture_input = texture_input
        self.shCoeff_input = shCoeff_input
        self.targetImage_input = targetImage_input
        self.extrinsics_input = extrinsics_input
        self.intrinsics_input = intrinsics_input

        self.nodeName = nodeName

        self.cudaRendererOperator = cuda_renderer_gpu(
            vertex_pos=self.vertexPos_input,
            vertex_color=self.vertexColor_input,
            texture=self.texture_input,
            sh_coeff=self.shCoeff_input,
            target_image=self.targetImage_input,
            extrinsics=self.extrinsics_input,
            intrinsics=self.intrinsics_input,
            faces=self.faces_attr,
            texture_coordinates=self.texCoords_attr,
            number_of_vertices=self.numberOfVertices_attr,
            number_of_cameras=self.numberOfCameras_attr,
            render_resolution_u=self.renderResolutionU_attr,
            render_resolution_v=self.renderResolutionV_attr,
            albedo_mode=self.albedoMode_attr,
            shading_mode=self.shadingMode_attr,
            image_filter_size=self.image_filter_size_attr,
            texture_filter_size=self.texture_filter_size_attr,
            compute_normal_map=self.compute_normal_map_attr,
            name=self.nodeName)

    def getBaryCentricBufferTF(self):
        return self.cudaRendererOperator[0]

    def getFaceBufferTF(self):
        return self.cudaRendererOperator[1]

    def getRenderBufferTF(self):
        return self.cudaRendererOperator[2]

    def getTargetBufferTF(self):
        return self.cudaRendererOperator[4]

    def getNormalMap(self):
        if self.compute_normal_map_attr:
            return self.cudaRendererOperator[5].reshape(torch.as_tensor(self.texture_input).shape)
        print('Requesting normal map but computation was not enabled!')
        return None

    def getModelMaskTF(self):
        face = self.cudaRendererOperator[1]
        return (face >= 0).to(torch.float32).unsqueeze(-1).expand(*face.shape, 3)

    def getBaryCentricBufferOpenCV(self, batchId, camId):
        # cv.cvtColor(RGB2BGR) of a 2-channel image is not defined in the reference either;
        # the channels are returned reversed.
        return self.cudaRendererOperator[0][batchId][camId].detach().cpu().numpy()[..., ::-1].copy()

    def getFaceBufferOpenCV(self, batchId, camId):
        faceImg = self.cudaRendererOperator[1][batchId][camId].detach().cpu().numpy().astype(np.float32) + 1.0
        return np.repeat(faceImg[..., None], 3, axis=-1)

    def getRenderBufferOpenCV(self, batchId, camId):
        return self.cudaRendererOperator[2][batchId][camId].detach().cpu().numpy()[..., ::-1].copy()

    def getNormalMapOpenCV(self, batchId):
        return self.cudaRendererOperator[5][batchId].detach().cpu().numpy()[..., ::-1].copy()
