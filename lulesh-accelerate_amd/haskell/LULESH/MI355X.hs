{-# LANGUAGE ForeignFunctionInterface #-}
-- |
-- Module      : LULESH.MI355X
-- Description : Haskell binding of the MI355X LULESH step
--
-- The drop-in for the Accelerate step of the LULESH driver: where
-- @simulate@ (src/Main.hs:90-109) advances the domain with
-- @$(CPU.runQ lulesh) mN0 v0@ (src/Main.hs:100-108), a driver built with
-- this module keeps the domain resident on the GPU and calls 'step'.  The
-- bindings follow the pattern of src/VisIt.hs:71-86 (@foreign import ccall@
-- of a plain C ABI on pointers and sizes); the ABI is
-- include/lulesh_mi355x.h and the library liblulesh_mi355x.so.
--
-- The module depends on base only.  Host arrays (for 'downloadState' and
-- 'uploadState') are borrowed for the duration of a call, so a driver can
-- pass the pointers of its Accelerate arrays exactly as src/VisIt.hs:63-68
-- obtains them with @ptrsOfArrayData@.
--
-- Every layout fact below (struct sizes and offsets, return codes, field
-- identifiers, argument lists) is checked against the C header by
-- tests/test_haskell_shim.py.
module LULESH.MI355X
  ( -- * Configuration and results
    Config(..)
  , Stats(..)
  , defaultConfig
    -- * Errors
  , LuleshError(..)
  , codeOk, codeVolumeError, codeQStopError, codeBadArg, codeHipError
    -- * Fields
  , Field(..)
  , fieldId
    -- * A resident domain
  , Domain
  , newDomain
  , newDomainWith
  , step
  , run
  , stats
  , downloadState
  , uploadState
  , getField
  , setField
  , setParameter
  , setFusedMultiplyAdd
  , symmetry
  ) where

import Control.Exception  ( Exception, throwIO )
import Control.Monad      ( when )
import Foreign
import Foreign.C.String
import Foreign.C.Types

-- | @lulesh_config@: the run options of lulesh-2.0.3's command line
-- (reference/C/lulesh-2.0.3/lulesh-util.cc:60-168).
data Config = Config
  { cfg_nx         :: !CInt     -- ^ -s: elements along each edge
  , cfg_ic_mode    :: !CInt     -- ^ 1: LULESH 1.0 ICs, 2: LULESH 2.0.3 ICs
  , cfg_num_reg    :: !CInt     -- ^ -r: regions (2.0 only)
  , cfg_balance    :: !CInt     -- ^ -b: region size balance (2.0 only)
  , cfg_cost       :: !CInt     -- ^ -c: extra cost of the costly regions
  , cfg_max_cycles :: !CInt     -- ^ -i: stop after this many cycles
  , cfg_stoptime   :: !CDouble  -- ^ stop when time >= stoptime
  } deriving (Eq, Show)

instance Storable Config where
  sizeOf _    = 32
  alignment _ = 8
  peek p = Config
    <$> peekByteOff p 0
    <*> peekByteOff p 4
    <*> peekByteOff p 8
    <*> peekByteOff p 12
    <*> peekByteOff p 16
    <*> peekByteOff p 20
    <*> peekByteOff p 24
  poke p c = do
    pokeByteOff p 0  (cfg_nx c)
    pokeByteOff p 4  (cfg_ic_mode c)
    pokeByteOff p 8  (cfg_num_reg c)
    pokeByteOff p 12 (cfg_balance c)
    pokeByteOff p 16 (cfg_cost c)
    pokeByteOff p 20 (cfg_max_cycles c)
    pokeByteOff p 24 (cfg_stoptime c)

-- | @lulesh_stats@: the scalars of the current state.
data Stats = Stats
  { st_time      :: !CDouble   -- ^ domain.time()
  , st_deltatime :: !CDouble   -- ^ domain.deltatime()
  , st_dtcourant :: !CDouble
  , st_dthydro   :: !CDouble
  , st_e_origin  :: !CDouble   -- ^ e(0), the "Final Origin Energy"
  , st_cycle     :: !CInt
  , st_done      :: !CInt      -- ^ time >= stoptime or cycle >= max_cycles
  , st_num_elem  :: !CInt
  , st_num_node  :: !CInt
  } deriving (Eq, Show)

instance Storable Stats where
  sizeOf _    = 56
  alignment _ = 8
  peek p = Stats
    <$> peekByteOff p 0
    <*> peekByteOff p 8
    <*> peekByteOff p 16
    <*> peekByteOff p 24
    <*> peekByteOff p 32
    <*> peekByteOff p 40
    <*> peekByteOff p 44
    <*> peekByteOff p 48
    <*> peekByteOff p 52
  poke p s = do
    pokeByteOff p 0  (st_time s)
    pokeByteOff p 8  (st_deltatime s)
    pokeByteOff p 16 (st_dtcourant s)
    pokeByteOff p 24 (st_dthydro s)
    pokeByteOff p 32 (st_e_origin s)
    pokeByteOff p 40 (st_cycle s)
    pokeByteOff p 44 (st_done s)
    pokeByteOff p 48 (st_num_elem s)
    pokeByteOff p 52 (st_num_node s)

-- | Return codes of every call; -1 and -2 mirror the reference's exits
-- (reference/C/LULESH-OMP.cc:82).
codeOk, codeVolumeError, codeQStopError, codeBadArg, codeHipError :: CInt
codeOk          = 0
codeVolumeError = -1
codeQStopError  = -2
codeBadArg      = -3
codeHipError    = -4

data LuleshError
  = VolumeError              -- ^ a volume went <= 0
  | QStopError               -- ^ q exceeded qstop
  | BadArgument String
  | HipError String
  deriving Show

instance Exception LuleshError

-- | The fields of 'getField' / 'setField' (node fields have num_node
-- values, element fields num_elem, in the reference's order).
data Field
  = F_X | F_Y | F_Z | F_XD | F_YD | F_ZD | F_NODAL_MASS
  | F_E | F_P | F_Q | F_QL | F_QQ | F_V | F_VOLO | F_DELV | F_VDOV | F_AREALG
  | F_SS | F_ELEM_MASS | F_VNEW
  | F_DELV_XI | F_DELV_ETA | F_DELV_ZETA | F_DELX_XI | F_DELX_ETA | F_DELX_ZETA
  deriving (Eq, Show, Enum, Bounded)

fieldId :: Field -> CInt
fieldId f = case f of
  F_X          -> 0
  F_Y          -> 1
  F_Z          -> 2
  F_XD         -> 3
  F_YD         -> 4
  F_ZD         -> 5
  F_NODAL_MASS -> 6
  F_E          -> 16
  F_P          -> 17
  F_Q          -> 18
  F_QL         -> 19
  F_QQ         -> 20
  F_V          -> 21
  F_VOLO       -> 22
  F_DELV       -> 23
  F_VDOV       -> 24
  F_AREALG     -> 25
  F_SS         -> 26
  F_ELEM_MASS  -> 27
  F_VNEW       -> 28
  F_DELV_XI    -> 29
  F_DELV_ETA   -> 30
  F_DELV_ZETA  -> 31
  F_DELX_XI    -> 32
  F_DELX_ETA   -> 33
  F_DELX_ZETA  -> 34

data LuleshDomain  -- the opaque lulesh_domain

-- | A domain resident on the GPU, freed by the garbage collector.
newtype Domain = Domain (ForeignPtr LuleshDomain)

foreign import ccall unsafe "lulesh_config_default"
  c_lulesh_config_default :: Ptr Config -> IO ()

foreign import ccall safe "lulesh_create"
  c_lulesh_create :: Ptr Config -> Ptr (Ptr LuleshDomain) -> IO CInt

foreign import ccall safe "&lulesh_destroy"
  p_lulesh_destroy :: FunPtr (Ptr LuleshDomain -> IO ())

foreign import ccall safe "lulesh_step"
  c_lulesh_step :: Ptr LuleshDomain -> CInt -> Ptr CInt -> IO CInt

foreign import ccall safe "lulesh_run"
  c_lulesh_run :: Ptr LuleshDomain -> Ptr CInt -> IO CInt

foreign import ccall safe "lulesh_get_stats"
  c_lulesh_get_stats :: Ptr LuleshDomain -> Ptr Stats -> IO CInt

foreign import ccall safe "lulesh_download_state"
  c_lulesh_download_state
    :: Ptr LuleshDomain
    -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble   -- x y z
    -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble   -- xd yd zd
    -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble   -- e p q
    -> Ptr CDouble -> Ptr CDouble                  -- v ss
    -> Ptr CDouble -> Ptr CDouble                  -- time deltatime
    -> IO CInt

foreign import ccall safe "lulesh_upload_state"
  c_lulesh_upload_state
    :: Ptr LuleshDomain
    -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble
    -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble
    -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble
    -> Ptr CDouble -> Ptr CDouble
    -> CDouble -> CDouble
    -> IO CInt

foreign import ccall safe "lulesh_get_field"
  c_lulesh_get_field :: Ptr LuleshDomain -> CInt -> Ptr CDouble -> IO CInt

foreign import ccall safe "lulesh_set_field"
  c_lulesh_set_field :: Ptr LuleshDomain -> CInt -> Ptr CDouble -> IO CInt

foreign import ccall safe "lulesh_set_parameter"
  c_lulesh_set_parameter :: Ptr LuleshDomain -> CString -> CDouble -> IO CInt

foreign import ccall safe "lulesh_set_fused_multiply_add"
  c_lulesh_set_fused_multiply_add :: Ptr LuleshDomain -> CInt -> IO CInt

foreign import ccall safe "lulesh_symmetry"
  c_lulesh_symmetry :: Ptr LuleshDomain -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble -> IO CInt

foreign import ccall unsafe "lulesh_last_error"
  c_lulesh_last_error :: Ptr LuleshDomain -> IO CString

-- | The reference defaults: nx 30, 2.0 ICs, 11 regions, balance 1, cost 1,
-- no cycle limit, stoptime 1e-2.
defaultConfig :: IO Config
defaultConfig = alloca $ \pc -> c_lulesh_config_default pc >> peek pc

-- | The cube of nx^3 elements with the ICs the Accelerate driver uses
-- (LULESH 1.0: unscaled origin energy, dt0 = 1e-7; src/Init.hs:27,
-- src/Main.hs:80) and at most maxSteps cycles.
newDomain :: Int -> Int -> IO Domain
newDomain nx maxSteps = do
  c <- defaultConfig
  newDomainWith c { cfg_nx = fromIntegral nx, cfg_ic_mode = 1
                  , cfg_max_cycles = fromIntegral maxSteps }

newDomainWith :: Config -> IO Domain
newDomainWith c = with c $ \pc -> alloca $ \pd -> do
  rc <- c_lulesh_create pc pd
  when (rc /= codeOk) $ throwIO (codeError rc "lulesh_create")
  d <- peek pd
  Domain <$> newForeignPtr p_lulesh_destroy d

codeError :: CInt -> String -> LuleshError
codeError rc what
  | rc == codeVolumeError = VolumeError
  | rc == codeQStopError  = QStopError
  | rc == codeBadArg      = BadArgument what
  | otherwise             = HipError what

check :: Ptr LuleshDomain -> CInt -> IO ()
check d rc = when (rc /= codeOk) $ do
  msg <- c_lulesh_last_error d >>= peekCString
  throwIO (codeError rc msg)

-- | One turn of @simulate@ (src/Main.hs:90-109): TimeIncrement then
-- LagrangeLeapFrog, up to n cycles; the cycles done.
step :: Domain -> Int -> IO Int
step (Domain fd) n = withForeignPtr fd $ \d -> alloca $ \pn -> do
  c_lulesh_step d (fromIntegral n) pn >>= check d
  fromIntegral <$> peek pn

-- | Every cycle to stoptime or the cycle limit; the cycles done.
run :: Domain -> IO Int
run (Domain fd) = withForeignPtr fd $ \d -> alloca $ \pn -> do
  c_lulesh_run d pn >>= check d
  fromIntegral <$> peek pn

stats :: Domain -> IO Stats
stats (Domain fd) = withForeignPtr fd $ \d -> alloca $ \ps -> do
  c_lulesh_get_stats d ps >>= check d
  peek ps

-- | The domain tuple of src/Domain.hs:11-21 into host arrays (x, y, z,
-- xd, yd, zd of num_node values; e, p, q, v, ss of num_elem values); the
-- time and time step.
downloadState :: Domain
              -> (Ptr CDouble, Ptr CDouble, Ptr CDouble)
              -> (Ptr CDouble, Ptr CDouble, Ptr CDouble)
              -> (Ptr CDouble, Ptr CDouble, Ptr CDouble, Ptr CDouble, Ptr CDouble)
              -> IO (CDouble, CDouble)
downloadState (Domain fd) (x, y, z) (xd, yd, zd) (e, p, q, v, ss) =
  withForeignPtr fd $ \d -> alloca $ \pt -> alloca $ \pdt -> do
    c_lulesh_download_state d x y z xd yd zd e p q v ss pt pdt >>= check d
    (,) <$> peek pt <*> peek pdt

uploadState :: Domain
            -> (Ptr CDouble, Ptr CDouble, Ptr CDouble)
            -> (Ptr CDouble, Ptr CDouble, Ptr CDouble)
            -> (Ptr CDouble, Ptr CDouble, Ptr CDouble, Ptr CDouble, Ptr CDouble)
            -> CDouble -> CDouble
            -> IO ()
uploadState (Domain fd) (x, y, z) (xd, yd, zd) (e, p, q, v, ss) t dt =
  withForeignPtr fd $ \d ->
    c_lulesh_upload_state d x y z xd yd zd e p q v ss t dt >>= check d

getField :: Domain -> Field -> Ptr CDouble -> IO ()
getField (Domain fd) f out = withForeignPtr fd $ \d ->
  c_lulesh_get_field d (fieldId f) out >>= check d

setField :: Domain -> Field -> Ptr CDouble -> IO ()
setField (Domain fd) f src = withForeignPtr fd $ \d ->
  c_lulesh_set_field d (fieldId f) src >>= check d

-- | A constant of the method by its reference name ("hgcoef", "qstop", ...).
setParameter :: Domain -> String -> Double -> IO ()
setParameter (Domain fd) name value = withForeignPtr fd $ \d ->
  withCString name $ \cn ->
    c_lulesh_set_parameter d cn (realToFrac value) >>= check d

-- | Float64 arithmetic with fused multiply-adds (True): about 10% faster,
-- results within the reference's tolerances rather than bit-identical.
setFusedMultiplyAdd :: Domain -> Bool -> IO ()
setFusedMultiplyAdd (Domain fd) on = withForeignPtr fd $ \d ->
  c_lulesh_set_fused_multiply_add d (if on then 1 else 0) >>= check d

-- | The plane-0 energy check of src/Main.hs:127-157: (maximum relative,
-- maximum absolute, total absolute difference).
symmetry :: Domain -> IO (Double, Double, Double)
symmetry (Domain fd) = withForeignPtr fd $ \d ->
  alloca $ \pa -> alloca $ \pt -> alloca $ \pr -> do
    c_lulesh_symmetry d pa pt pr >>= check d
    a <- peek pa
    t <- peek pt
    r <- peek pr
    return (realToFrac r, realToFrac a, realToFrac t)
