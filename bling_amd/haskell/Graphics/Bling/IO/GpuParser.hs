-- | Parser side of the @gpu@ renderer: the renderer statement and the
-- capture of the flat scene description while the reference's parser
-- builds its primitives.
--
-- A maintainer adds three things to the reference's parser:
--
--   * a field @gpuScene :: IORef GpuSceneBuilder@ in PState
--     (IO/ParserCore.hs:45-58), created with 'newGpuScene' where the
--     initial state is made;
--   * one case in pRenderer (IO/RendererParser.hs:23-54):
--     @"gpu" -> pGpuRenderer (gpuScene s)@;
--   * one call next to each constructor the parser already calls:
--     'captureMaterial' beside the material parsers
--     (IO/MaterialParser.hs:30-104), 'captureMesh' beside mkTriangleMesh and
--     mkHeightMap (IO/PrimitiveParser.hs:28-139), 'captureShape' beside
--     mkGeom, 'captureInfinite' beside mkInfiniteAreaLight
--     (IO/LightParser.hs:29-33), 'captureView' where the camera and filter
--     are set (IO/CameraParser.hs:30-38, IO/ParserCore.hs).
--
-- The capture functions take the values those constructors receive, so
-- the description is the one the reference renders.  They need the 16
-- bands of a Spectrum and a Transform's matrices, which the reference keeps
-- behind its export lists: 'spectrumBands' and 'transformRows' are the two
-- accessors a maintainer exports (Spectrum.hs:46, Transform.hs:123-127).
module Graphics.Bling.IO.GpuParser (
   newGpuScene, pGpuRenderer, GpuMaterial(..), captureMaterial, captureMesh, captureShape,
   captureInfinite, captureView
   ) where

import Control.Monad.IO.Class (liftIO)
import Data.IORef
import qualified Data.Vector as V
import qualified Data.Vector.Storable as VS
import Foreign (nullPtr)
import Text.ParserCombinators.Parsec

import Graphics.Bling.Filter (Filter)
import Graphics.Bling.GPU
import Graphics.Bling.GPU.FFI
import Graphics.Bling.IO.ParserCore
import Graphics.Bling.Math
import qualified Graphics.Bling.Shape as S
import Graphics.Bling.Rendering (AnyRenderer, mkAnyRenderer)
import Graphics.Bling.Spectrum (Spectrum, spectrumBands)
import Graphics.Bling.Transform (Transform, transNormal, transPoint, transformRows)

newGpuScene :: IO (IORef GpuSceneBuilder)
newGpuScene = newIORef emptyGpuScene

-- | @gpu sampled { sampler { stratified nu nv } integrator { path maxDepth md
-- sampleDepth sd } }@ or with @integrator { directLighting maxDepth md }@:
-- the syntax of the sampler renderer (IO/RendererParser.hs:56-73,
-- IO/IntegratorParser.hs:15-50); the GPU path samples stratified only.
pGpuRenderer :: IORef GpuSceneBuilder -> JobParser AnyRenderer
pGpuRenderer ref = (flip namedBlock) "sampled" $ do
   (nu, nv) <- (flip namedBlock) "sampler" $ pString >>= \t -> case t of
      "stratified" -> do
         nx <- integ
         ny <- integ
         return (nx, ny)
      _ -> fail $ "the gpu renderer samples stratified only, not " ++ t
   job <- (flip namedBlock) "integrator" $ pString >>= \t -> case t of
      "path" -> do
         md <- namedInt "maxDepth"
         sd <- namedInt "sampleDepth"
         return defaultGpuJob { gjMaxDepth = md, gjSampleDepth = sd }
      "directLighting" -> do
         md <- namedInt "maxDepth"
         return defaultGpuJob { gjMaxDepth = md, gjSampleDepth = 0, gjDirect = True }
      _ -> fail $ "the gpu renderer has no integrator " ++ t
   return $ mkAnyRenderer $ mkGpuRenderer ref job { gjNu = nu, gjNv = nv }

-- | The materials on the GPU path, with the values their parsers read
-- (Material.hs:32-127); textured slots are refused by the caller unless
-- they are constants or a checker / graphPaper selector over constants.
data GpuMaterial
   = GMatte Spectrum Float                 -- ^ kd, sigma
   | GPlastic Spectrum Spectrum Float      -- ^ kd, ks, roughness
   | GMetal Spectrum Spectrum Float        -- ^ eta, k, roughness
   | GGlass Spectrum Spectrum Float        -- ^ kr, kt, ior
   | GMirror Spectrum                      -- ^ kr
   | GTransMatte Spectrum Spectrum Float   -- ^ kr, kt, sigma
   | GSubstrate Spectrum Spectrum Spectrum Float Float Float  -- ^ kd, ks, ka, urough, vrough, depth
   | GBlack

noTexture :: Texture
noTexture = Texture texNone mapUV 0 0 (replicate 12 0) [replicate 16 0, replicate 16 0]

-- | Records a material; returns its descriptor id.
captureMaterial :: IORef GpuSceneBuilder -> GpuMaterial -> Maybe Texture -> JobParser Int32
captureMaterial ref gm tex = liftIO $ atomicModifyIORef' ref (\b -> let (i, b') = gpuAddMaterial m b in (b', i))
   where
      m = case gm of
         GMatte kd s -> rec matMatte [s] [kd]
         GPlastic kd ks r -> rec matPlastic [r] [kd, ks]
         GMetal e k r -> rec matMetal [r] [e, k]
         GGlass kr kt ior -> rec matGlass [ior] [kr, kt]
         GMirror kr -> rec matMirror [] [kr]
         GTransMatte kr kt s -> rec matTransMatte [s] [kr, kt]
         GSubstrate kd ks ka u v d -> rec matSubstrate [u, v, d] [kd, ks, ka]
         GBlack -> rec matBlack [] []
      rec t sc sp = Material t 0 (take 4 (sc ++ repeat 0)) (take 3 (map spectrumBands sp ++ repeat (replicate 16 0)))
                             (maybe noTexture id tex)

-- | mkTriangleMesh o2w mat p i n uv (Primitive/TriangleMesh.hs:39-59): the
-- vertices and normals go to world space with the same transPoint /
-- transNormal calls, so the descriptor holds the values the reference
-- intersects.
captureMesh :: IORef GpuSceneBuilder -> Transform -> Int32 -> V.Vector Point -> V.Vector Int
            -> Maybe (V.Vector Normal) -> Maybe (V.Vector (Float, Float)) -> JobParser ()
captureMesh ref o2w mat ps is ns uvs = liftIO $ modifyIORef' ref (gpuAddMesh mat pv iv nv uvv) where
   xyz v = [vx v, vy v, vz v]
   pv = VS.fromList $ concatMap (xyz . transPoint o2w) (V.toList ps)
   iv = VS.fromList $ map fromIntegral (V.toList is)
   nv = fmap (\n -> VS.fromList $ concatMap (xyz . transNormal o2w) (V.toList n)) ns
   uvv = fmap (\u -> VS.fromList $ concatMap (\(a, b) -> [a, b]) (V.toList u)) uvs

-- | mkGeom o2w _ mat emission shape gid (Primitive/Geometry.hs:14-37).
captureShape :: IORef GpuSceneBuilder -> Transform -> Int32 -> Maybe Spectrum -> S.Shape -> JobParser ()
captureShape ref o2w mat em s = liftIO $ modifyIORef' ref $ \b ->
   let (o2wRows, w2oRows) = transformRows o2w
       (ty, ps) = case s of
          S.Box p0 p1 -> (shapeBox, [vx p0, vy p0, vz p0, vx p1, vy p1, vz p1])
          S.Cylinder r z0 z1 phi -> (shapeCylinder, [r, z0, z1, phi])
          S.Disk h r ri phi -> (shapeDisk, [h, r, ri, phi])
          S.Quad sx sy -> (shapeQuad, [sx, sy])
          S.Sphere r -> (shapeSphere, [r])
       sh = Shape ty mat (maybe 0 (const 1) em) (gpuNextLightId b) (take 6 (ps ++ repeat 0))
                  (concat o2wRows) (concat w2oRows) (maybe (replicate 16 0) spectrumBands em)
   in gpuAddShape sh b

-- | mkInfiniteAreaLight rmap t (Light.hs:72-82) for a map of w x h texels,
-- each 16 bands, row-major from the top (the texel lookup of Texture.hs:94-101).
captureInfinite :: IORef GpuSceneBuilder -> Transform -> Int -> Int -> [Spectrum] -> JobParser ()
captureInfinite ref w2l w h texels = liftIO $ modifyIORef' ref (gpuAddInfiniteLight l m) where
   (w2lRows, l2wRows) = transformRows w2l
   l = Light lightInfinite (fromIntegral w) (fromIntegral h) 0 nullPtr (concat w2lRows) (concat l2wRows)
   m = VS.fromList (concatMap spectrumBands texels)

-- | The camera after mkProjective (Camera.hs:108-147) and the film
-- (imageSize and filter, Filter.hs:12-96).
captureView :: IORef GpuSceneBuilder -> Transform -> Transform -> Float -> Float -> Film -> Filter -> JobParser ()
captureView ref c2w r2c lensR focalD film flt = liftIO $ modifyIORef' ref (gpuSetFilm film flt . gpuSetCamera cam) where
   cam = Camera cameraPerspective 0 (concat (fst (transformRows c2w))) (concat (fst (transformRows r2c)))
                lensR focalD 0 0
